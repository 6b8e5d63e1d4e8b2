// hml_dense.cpp -- dispatch of the dense SGEMM kernels (plain FP32 MFMA and
// split-precision bf16x3 / bf16x6 / scaled f16x3 MFMA).  No reference counterpart: the
// reference routes every shape through its skinny kernels
// (src/hml_sgemm.cu:120-211), whose 96 x 96 basic tile is its only dense
// path.
#include <algorithm>

#include "hml_dense.h"
#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_select.h"

namespace hml {

namespace {

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// bf16x3 and scaled f16x3 keep 2 planes per operand, bf16x6 3, FP32 none.
int split_terms(int prec) { return prec == 1 || prec == 3 ? 2 : prec == 2 ? 3 : 0; }

struct DensePlan {
  int64_t Mp, Np, Kp;
  size_t a_bytes, b_bytes;
  size_t rmax_bytes;  // f16x3: row maxima of A (Mp) then of B^T (Np)
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
  d.rmax_bytes = prec == 3 ? align256((size_t)(d.Mp + d.Np) * sizeof(uint32_t)) : 0;
  return d;
}

}  // namespace

size_t dense_workspace(int prec, int64_t M, int64_t N, int64_t K) {
  const DensePlan d = plan_dense(prec, M, N, K);
  return d.a_bytes + d.b_bytes + d.rmax_bytes;
}

// C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n) with A(m,k) =
// A[m*sAm + k*sAk], B(k,n) = B[k*sBk + n*sBn], C(m,n) = C[m*sCm + n*sCn].
hmlStatus run_dense(int prec, const float *A, int64_t sAm, int64_t sAk,
                    const float *B, int64_t sBk, int64_t sBn, float *C,
                    int64_t sCm, int64_t sCn, int64_t M, int64_t N, int64_t K,
                    float alpha, float beta, int read_c, void *ws,
                    size_t ws_bytes, hipStream_t st) {
  const DensePlan d = plan_dense(prec, M, N, K);
  const size_t need = d.a_bytes + d.b_bytes + d.rmax_bytes;
  WsLease lease;  // held until the kernels below are queued
  if (!ws) {
    hmlStatus s = library_workspace(need, st, &lease);
    if (s != HML_STATUS_SUCCESS) return s;
    ws = lease.ptr();
  } else if (ws_bytes < need) {
    return fail(HML_STATUS_WORKSPACE_TOO_SMALL, "workspace %zu < %zu bytes", ws_bytes, need);
  }
  const int S = split_terms(prec);
  void *Ap = ws;
  void *Bp = (char *)ws + d.a_bytes;
  uint32_t *rmA = nullptr, *rmB = nullptr;
  hipError_t e;
  if (prec == 3) {
    // scaled fp16 split: zero the row maxima (a fill kernel, so the sequence
    // can be captured), gather them, then split with the row scales
    rmA = (uint32_t *)((char *)ws + d.a_bytes + d.b_bytes);
    rmB = rmA + d.Mp;
    e = launch_fill((float *)rmA, (size_t)(d.Mp + d.Np), 0.f, st);
    if (e == hipSuccess) e = launch_rowmax(A, sAm, sAk, M, K, rmA, st);
    if (e == hipSuccess) e = launch_rowmax(B, sBn, sBk, N, K, rmB, st);
    if (e == hipSuccess) e = launch_split_f16(A, sAm, sAk, M, K, d.Mp, d.Kp, rmA, Ap, st);
    if (e == hipSuccess) e = launch_split_f16(B, sBn, sBk, N, K, d.Np, d.Kp, rmB, Bp, st);
  } else {
    e = launch_split(A, sAm, sAk, M, K, d.Mp, d.Kp, S, Ap, st);
    if (e == hipSuccess) e = launch_split(B, sBn, sBk, N, K, d.Np, d.Kp, S, Bp, st);
  }
  if (e != hipSuccess) return from_hip(e, "split/pack pass");
  DenseParams p{};
  p.Ap = Ap; p.Bp = Bp; p.C = C; p.sCm = sCm; p.sCn = sCn;
  p.M = M; p.N = N; p.Mp = d.Mp; p.Np = d.Np; p.Kp = d.Kp;
  p.alpha = alpha; p.beta = beta; p.read_c = read_c;
  p.rmaxA = rmA; p.rmaxB = rmB;
  e = launch_dense(prec == 3 ? 4 : S, p, st);
  return from_hip(e, "dense kernel");
}

}  // namespace hml

using namespace hml;

extern "C" {

size_t hmlSgemmDenseWorkspaceSize(int32_t precision, int64_t M, int64_t N,
                                  int64_t K) {
  if (M <= 0 || N <= 0 || K <= 0 || precision < 0 || precision > 3) return 0;
  return dense_workspace(precision, M, N, K);
}

hmlStatus hmlSgemmDense(int32_t precision, hmlOrder order, hmlOp transA,
                        hmlOp transB, int64_t M, int64_t N, int64_t K,
                        float alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, float beta, float *C,
                        int64_t ldc, uint32_t flags, void *workspace,
                        size_t ws_bytes, void *stream) {
  clear_error();
  if (precision < 0 || precision > 3)
    return fail(HML_STATUS_INVALID_VALUE,
                "precision must be 0 (FP32), 1 (bf16x3), 2 (bf16x6) or 3 (f16x3)");
  static const int32_t fam[4] = {HML_SGEMM_FAMILY_DENSE, HML_SGEMM_FAMILY_BF16X3,
                                 HML_SGEMM_FAMILY_BF16X6, HML_SGEMM_FAMILY_F16X3};
  hmlSgemmVariant v{fam[precision], 1, 1, 1, 4};
  return hmlSgemmExVariant(order, transA, transB, M, N, K, alpha, A, lda, B, ldb,
                           beta, C, ldc, flags, &v, workspace, ws_bytes, stream);
}

}  // extern "C"
