// hml_kmeans.cpp -- hmlKmeansAssign: argument checks and launch of the
// nearest-centroid kernels (hml_k_kmeans.hip): the direct form for few
// clusters x dims, the MFMA-screened, exactly refined form above that.
//
// Reference flow being replaced: hmlKmeansGpu picks one of its assign
// kernels per (numDims, numClusts) from a config table and launches it
// (src/hml_kmeans_gpu.cu:77-87, 136-137); here one kernel shape per
// (dims / 16, centroid tiles per pass) covers every size.
#include "hml_internal.h"
#include "hml_kmeans.h"
#include "hml_runtime.h"

using namespace hml;

extern "C" hmlStatus hmlKmeansAssign(uint32_t *assign, const float *rows, int64_t numRows,
                                     int32_t numDims, const float *ctrds, int32_t numClusts,
                                     uint32_t flags, void *stream) {
  clear_error();
  if (numRows < 0 || numDims < 1 || numDims > 64 || numClusts < 1)
    return fail(HML_STATUS_INVALID_VALUE,
                "kmeans assign: numRows %lld, numDims %d (1..64), numClusts %d (>= 1)",
                (long long)numRows, numDims, numClusts);
  if (flags & ~(HML_KMEANS_SEPARATE_ROUNDING | HML_KMEANS_KERNEL_SCREENED | HML_KMEANS_KERNEL_DIRECT) ||
      ((flags & HML_KMEANS_KERNEL_SCREENED) && (flags & HML_KMEANS_KERNEL_DIRECT)))
    return fail(HML_STATUS_INVALID_VALUE, "kmeans assign: bad flags 0x%x", flags);
  if (numRows == 0) return HML_STATUS_SUCCESS;
  if (!assign || !rows || !ctrds)
    return fail(HML_STATUS_INVALID_VALUE, "kmeans assign: null pointer");
  if (numRows > (int64_t)0xffffffff)
    return fail(HML_STATUS_INVALID_VALUE, "kmeans assign: more than 2^32 - 1 rows");
  const int mode = (flags & HML_KMEANS_KERNEL_SCREENED) ? 1 : (flags & HML_KMEANS_KERNEL_DIRECT) ? 2 : 0;
  if (mode == 2 && numDims > kKmDirectMaxDims)
    return fail(HML_STATUS_UNSUPPORTED, "kmeans assign: the direct kernel takes at most %d dims",
                kKmDirectMaxDims);
  const KmeansPlan pl = plan_kmeans(numRows, numDims, numClusts, mode);
  if (pl.direct ? pl.KC < 1 : pl.KC < 16 * pl.nt)
    return fail(HML_STATUS_UNSUPPORTED, "kmeans assign: no LDS chunk for %d dims", numDims);
  hipError_t e = launch_kmeans_assign(pl, assign, rows, numRows, numDims, ctrds, numClusts,
                                      !(flags & HML_KMEANS_SEPARATE_ROUNDING),
                                      (hipStream_t)stream);
  return e == hipSuccess ? HML_STATUS_SUCCESS : from_hip(e, "kmeans assign launch");
}
