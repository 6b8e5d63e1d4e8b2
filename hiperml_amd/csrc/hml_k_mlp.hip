// hml_k_mlp.hip -- training and inference kernels of the learned kernel
// selector: a small FP32 MLP (inputs -> hidden ReLU layers -> class
// logits), trained with a masked soft-target cross-entropy and Adam.  No
// reference counterpart (the reference selects kernels from a static
// table, src/hml_sgemm_kernel.cu:189-237).
//
// The network is tiny (a few thousand weights, batches of a few hundred
// rows), so each kernel maps one thread to one output element and reduces
// over the batch or the fan-in serially in a fixed order: results are
// deterministic run to run and match a sequential CPU restatement up to
// FP32 rounding.  Weights are stored [in][out] (row i holds the fan-out of
// input i).
#include "hml_internal.h"
#include "hml_mlp.h"

namespace hml {
namespace {

// Y[b][o] = act(bias[o] + sum_i X[b][i] W[i][o])
__global__ void k_linear_fwd(const float *__restrict__ X, const float *__restrict__ W,
                             const float *__restrict__ bias, float *__restrict__ Y,
                             int B, int I, int O, int relu) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * O) return;
  const int b = idx / O, o = idx - b * O;
  float s = bias[o];
  for (int i = 0; i < I; ++i) s = fmaf(X[b * I + i], W[i * O + o], s);
  Y[idx] = relu ? fmaxf(s, 0.f) : s;
}

// Masked softmax cross-entropy against soft targets T (rows sum to 1 over
// the mask).  loss_row[b] = -sum_o T log p; dZ = (p - T) / B on allowed
// classes, 0 elsewhere.  probs receives p (0 on masked classes).
__global__ void k_softmax_xent(const float *__restrict__ Z, const float *__restrict__ T,
                               const float *__restrict__ mask, float *__restrict__ probs,
                               float *__restrict__ dZ, float *__restrict__ loss_row,
                               int B, int O, float inv_batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const float *z = Z + (size_t)b * O;
  const float *m = mask + (size_t)b * O;
  float mx = -3.0e38f;
  for (int o = 0; o < O; ++o)
    if (m[o] > 0.f) mx = fmaxf(mx, z[o]);
  float den = 0.f;
  for (int o = 0; o < O; ++o)
    if (m[o] > 0.f) den += expf(z[o] - mx);
  const float lden = logf(den);
  float loss = 0.f;
  for (int o = 0; o < O; ++o) {
    float p = 0.f, g = 0.f;
    if (m[o] > 0.f) {
      const float lp = z[o] - mx - lden;
      p = expf(lp);
      if (T) {
        const float t = T[(size_t)b * O + o];
        loss -= t * lp;
        g = (p - t) * inv_batch;
      }
    }
    if (probs) probs[(size_t)b * O + o] = p;
    if (dZ) dZ[(size_t)b * O + o] = g;
  }
  if (loss_row) loss_row[b] = loss;
}

// dX[b][i] = sum_o dZ[b][o] W[i][o], then masked by the ReLU of the layer
// that produced X (Xact > 0), giving the pre-activation gradient of the
// previous layer.
__global__ void k_linear_bwd_data(const float *__restrict__ dZ, const float *__restrict__ W,
                                  const float *__restrict__ Xact, float *__restrict__ dX,
                                  int B, int I, int O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * I) return;
  const int b = idx / I, i = idx - b * I;
  float s = 0.f;
  for (int o = 0; o < O; ++o) s = fmaf(dZ[b * O + o], W[i * O + o], s);
  dX[idx] = Xact[idx] > 0.f ? s : 0.f;
}

// dW[i][o] = sum_b X[b][i] dZ[b][o] (i < I), db[o] = sum_b dZ[b][o] (row I)
__global__ void k_linear_bwd_weight(const float *__restrict__ X, const float *__restrict__ dZ,
                                    float *__restrict__ dW, float *__restrict__ db,
                                    int B, int I, int O) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (I + 1) * O) return;
  const int i = idx / O, o = idx - i * O;
  float s = 0.f;
  if (i < I) {
    for (int b = 0; b < B; ++b) s = fmaf(X[b * I + i], dZ[b * O + o], s);
    dW[i * O + o] = s;
  } else {
    for (int b = 0; b < B; ++b) s += dZ[b * O + o];
    db[o] = s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size.
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, int n,
                       float lr_t, float b1, float b2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// sum of n values in index order (one thread)
__global__ void k_sum_serial(const float *__restrict__ x, int n, float *__restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += x[i];
    *out = s;
  }
}

// dst row r = src row idx[r] (minibatch gather of a shuffled epoch)
__global__ void k_gather_rows(const float *__restrict__ src, const int32_t *__restrict__ idx,
                              float *__restrict__ dst, int rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * cols) return;
  const int64_t r = i / cols, c = i - r * cols;
  dst[i] = src[(int64_t)idx[r] * cols + c];
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

hipError_t mlp_gather_rows(const float *src, const int32_t *idx, float *dst,
                           int rows, int cols, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_rows, grid_for((int64_t)rows * cols), dim3(256),
                     0, st, src, idx, dst, rows, cols);
  return hipGetLastError();
}

hipError_t mlp_linear_fwd(const float *X, const float *W, const float *bias,
                          float *Y, int B, int I, int O, int relu,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_linear_fwd, grid_for((int64_t)B * O), dim3(256), 0, st,
                     X, W, bias, Y, B, I, O, relu);
  return hipGetLastError();
}

hipError_t mlp_softmax_xent(const float *Z, const float *T, const float *mask,
                            float *probs, float *dZ, float *loss_row, int B,
                            int O, hipStream_t st) {
  hipLaunchKernelGGL(k_softmax_xent, grid_for(B), dim3(256), 0, st, Z, T, mask,
                     probs, dZ, loss_row, B, O, 1.0f / (float)B);
  return hipGetLastError();
}

hipError_t mlp_linear_bwd_data(const float *dZ, const float *W,
                               const float *Xact, float *dX, int B, int I,
                               int O, hipStream_t st) {
  hipLaunchKernelGGL(k_linear_bwd_data, grid_for((int64_t)B * I), dim3(256), 0,
                     st, dZ, W, Xact, dX, B, I, O);
  return hipGetLastError();
}

hipError_t mlp_linear_bwd_weight(const float *X, const float *dZ, float *dW,
                                 float *db, int B, int I, int O,
                                 hipStream_t st) {
  hipLaunchKernelGGL(k_linear_bwd_weight, grid_for((int64_t)(I + 1) * O),
                     dim3(256), 0, st, X, dZ, dW, db, B, I, O);
  return hipGetLastError();
}

hipError_t mlp_adam(float *w, const float *g, float *m, float *v, int n,
                    float lr_t, float b1, float b2, float eps,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_adam, grid_for(n), dim3(256), 0, st, w, g, m, v, n,
                     lr_t, b1, b2, eps);
  return hipGetLastError();
}

hipError_t mlp_sum(const float *x, int n, float *out, hipStream_t st) {
  hipLaunchKernelGGL(k_sum_serial, dim3(1), dim3(64), 0, st, x, n, out);
  return hipGetLastError();
}

}  // namespace hml
