// hml_mlp.h -- learned kernel selector: MLP kernels and class space
// (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hml_ext.h"

namespace hml {

hipError_t mlp_linear_fwd(const float *X, const float *W, const float *bias,
                          float *Y, int B, int I, int O, int relu,
                          hipStream_t st);
hipError_t mlp_softmax_xent(const float *Z, const float *T, const float *mask,
                            float *probs, float *dZ, float *loss_row, int B,
                            int O, hipStream_t st);
hipError_t mlp_linear_bwd_data(const float *dZ, const float *W,
                               const float *Xact, float *dX, int B, int I,
                               int O, hipStream_t st);
hipError_t mlp_linear_bwd_weight(const float *X, const float *dZ, float *dW,
                                 float *db, int B, int I, int O,
                                 hipStream_t st);
hipError_t mlp_adam(float *w, const float *g, float *m, float *v, int n,
                    float lr_t, float b1, float b2, float eps, hipStream_t st);
hipError_t mlp_sum(const float *x, int n, float *out, hipStream_t st);
hipError_t mlp_gather_rows(const float *src, const int32_t *idx, float *dst,
                           int rows, int cols, hipStream_t st);

// Selector class space.  Classes 0..83: skinny MFMA variants (rt index 0..2
// -> 1/2/4 row tiles, column mode 0..1 -> full / half the column tiles that
// cover N (DOT: at most 4), ksplit 2^0..2^6, wpb 4 or 8); 84: dense FP32
// tiles; 85: the GEMV kernels (heuristic parameters); 86: the multi-vector
// GEMV kernels for N = 2, 4, 8 with A contiguous along K (the SGEMV
// heuristic's parameters for the same rows); 87..110: GEMV DOT geometries
// (lanes per row 8/16/32/64 x rows per lane group 1/2/4 x 2 or 8 waves per
// workgroup); 111..118: GEMV AXPY geometries (lanes per column chunk
// 8/16/32/64 x 4 or 8 waves); a GEMV class's row split follows the SGEMV
// split rule for its geometry.  The family of a skinny class follows the
// problem's form (dot or axpy).
constexpr int kSelFeatures = 6;
constexpr int kSelGemvDotFirst = 87;
constexpr int kSelGemvAxpyFirst = 111;
constexpr int kSelClasses = 119;

}  // namespace hml
