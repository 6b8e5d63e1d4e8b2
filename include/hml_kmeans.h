/*
 * hml_kmeans.h -- k-means nearest-centroid assignment on the MI355X.
 *
 * The reference runs its assign step as one of three CUDA kernels picked
 * per (numDims, numClusts) from a table (include/hml_kmeans_kernel.h:17-21,
 * include/hml_kmeans_kernel_template.h:18-88, 90-316, 323-367), launched
 * from hmlKmeansGpu (src/hml_kmeans_gpu.cu:136-137); its CPU path is
 * hmlKmeansAssign (src/hml_kmeans_cpu.cu:67-95).  Both compute, per row and
 * per centroid k ascending, dist = sum over dims ascending of (x - c)^2 in
 * float and keep the first k whose dist is strictly below every earlier one
 * (starting from FLT_MAX).  This entry point returns exactly those indices.
 *
 * Pointers are DEVICE pointers; `stream` is a hipStream_t passed as void*
 * (NULL = the null stream); the call is asynchronous.
 */
#ifndef HML_KMEANS_H_INCLUDED_
#define HML_KMEANS_H_INCLUDED_

#include <stdint.h>

#include "hml_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the distance sum `dist += delta * delta`:
 * default      one fused multiply-add per dimension, as the reference's
 *              GPU kernels are built (nvcc contracts it by default);
 * SEPARATE     product and sum rounded separately, as the reference's CPU
 *              path runs on baseline x86-64. */
#define HML_KMEANS_SEPARATE_ROUNDING 0x1u

/* Kernel choice; the indices are the same either way.  By default up to 8
 * dims, or numClusts x numDims < 1024 up to 32 dims, run the direct kernel
 * (one row per lane, the reference loop), larger problems the screened
 * kernel (x.c on the matrix cores, exact refinement of the candidates).  These force one;
 * the direct kernel takes at most 32 dims (HML_STATUS_UNSUPPORTED above). */
#define HML_KMEANS_KERNEL_SCREENED 0x2u
#define HML_KMEANS_KERNEL_DIRECT   0x4u

/* assign[r] = the nearest centroid of row r (r < numRows).
 *   rows   numRows x numDims floats, row-major (the reference's pRows)
 *   ctrds  numClusts x numDims floats, row-major (pCtrds)
 *   assign numRows uint32 (pAsmnts)
 * numDims 1..64 (cHmlKmeansMaxDims, include/hml_kmeans_kernel.h:13),
 * numClusts >= 1, numRows >= 0.  A row whose distances are all NaN or
 * >= FLT_MAX gets numClusts (the reference leaves such rows undefined on
 * the GPU).  Replaces the assign-kernel launch at
 * src/hml_kmeans_gpu.cu:136-137. */
hmlStatus hmlKmeansAssign(uint32_t *assign, const float *rows, int64_t numRows,
                          int32_t numDims, const float *ctrds, int32_t numClusts,
                          uint32_t flags, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* HML_KMEANS_H_INCLUDED_ */
