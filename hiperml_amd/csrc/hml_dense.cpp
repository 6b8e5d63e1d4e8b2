// hml_dense.cpp -- dispatch of the dense SGEMM kernels (plain FP32 MFMA and
// split-precision bf16x3 / bf16x6 MFMA).  No reference counterpart: the
// reference routes every shape through its skinny kernels
// (src/hml_sgemm.cu:120-211), whose 96 x 96 basic tile is its only dense
// path.
#include <algorithm>

#include "hml_dense.h"
#include "hml_runtime.h"
#include "hml_select.h"

namespace hml {

namespace {

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

int split_terms(int prec) { return prec == 1 ? 2 : prec == 2 ? 3 : 0; }

struct DensePlan {
  int64_t Mp, Np, Kp;
  size_t a_bytes, b_bytes;
};

DensePlan plan_dense(int prec, int64_t M, int64_t N, int64_t K) {
  DensePlan d;
  const int S = split_terms(prec);
  d.Mp = round_up(std::max<int64_t>(M, 1), S ? kSplitTileM : kDenseTileM);
  d.Np = round_up(std::max<int64_t>(N, 1), S ? kSplitTileN : kDenseTileN);
  d.Kp = round_up(std::max<int64_t>(K, 1), kDenseTileK);
  const size_t elem = S ? 2 : 4;
  const int planes = S ? S : 1;
  d.a_bytes = align256((size_t)planes * d.Mp * d.Kp * elem);
  d.b_bytes = align256((size_t)planes * d.Np * d.Kp * elem);
  return d;
}

}  // namespace

size_t dense_workspace(int prec, int64_t M, int64_t N, int64_t K) {
  const DensePlan d = plan_dense(prec, M, N, K);
  return d.a_bytes + d.b_bytes;
}

// C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n) with A(m,k) =
// A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn], C(m,n) = C[m*sCm + n*sCn].
hmlStatus run_dense(int prec, const float *A, int64_t sAm, int64_t sAk,
                    const float *B, int64_t sBk, int64_t sBn, float *C,
                    int64_t sCm, int64_t sCn, int64_t M, int64_t N, int64_t K,
                    float alpha, float beta, int read_c, void *ws,
                    size_t ws_bytes, hipStream_t st) {
  const DensePlan d = plan_dense(prec, M, N, K);
  const size_t need = d.a_bytes + d.b_bytes;
  if (!ws) {
    hmlStatus s = library_workspace(need, &ws);
    if (s != HML_STATUS_SUCCESS) return s;
  } else if (ws_bytes < need) {
    return fail(HML_STATUS_WORKSPACE_TOO_SMALL, "workspace %zu < %zu bytes", ws_bytes, need);
  }
  const int S = split_terms(prec);
  void *Ap = ws;
  void *Bp = (char *)ws + d.a_bytes;
  hipError_t e = launch_split(A, sAm, sAk, M, K, d.Mp, d.Kp, S, Ap, st);
  if (e == hipSuccess) e = launch_split(B, sBn, sBk, N, K, d.Np, d.Kp, S, Bp, st);
  if (e != hipSuccess) return from_hip(e, "split/pack pass");
  DenseParams p{};
  p.Ap = Ap; p.Bp = Bp; p.C = C; p.sCm = sCm; p.sCn = sCn;
  p.M = M; p.N = N; p.Mp = d.Mp; p.Np = d.Np; p.Kp = d.Kp;
  p.alpha = alpha; p.beta = beta; p.read_c = read_c;
  e = launch_dense(S, p, st);
  return from_hip(e, "dense kernel");
}

}  // namespace hml

using namespace hml;

extern "C" {

size_t hmlSgemmDenseWorkspaceSize(int32_t precision, int64_t M, int64_t N,
                                  int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0 || precision < 0 || precision > 2) return 0;
  return dense_workspace(precision, M, N, K);
}

hmlStatus hmlSgemmDense(int32_t precision, hmlOrder order, hmlOp transA,
                        hmlOp transB, int64_t M, int64_t N, int64_t K,
                        float alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, float beta, float *C,
                        int64_t ldc, uint32_t flags, void *workspace,
                        size_t ws_bytes, void *stream) {
  clear_error();
  if (precision < 0 || precision > 2)
    return fail(HML_STATUS_INVALID_VALUE, "precision must be 0 (FP32), 1 (bf16x3) or 2 (bf16x6)");
  hmlSgemmVariant v{precision == 0 ? HML_SGEMM_FAMILY_DENSE
                                   : precision == 1 ? HML_SGEMM_FAMILY_BF16X3
                                                    : HML_SGEMM_FAMILY_BF16X6,
                    1, 1, 1, 4};
  return hmlSgemmExVariant(order, transA, transB, M, N, K, alpha, A, lda, B, ldb,
                           beta, C, ldc, flags, &v, workspace, ws_bytes, stream);
}

}  // extern "C"
