// hml_gemm.cpp -- SGEMM dispatch: the reference-compatible hmlSgemm /
// hmlSgemmBasic (host pointers, exit on error) and the extended
// device-pointer entry points.
//
// Reference flow being replaced (src/hml_sgemm.cu:120-211): lazy repo init,
// lazy online autotune, cudaMalloc + H2D + texture bind per call, table
// select (src/hml_sgemm_kernel.cu:189-237), launch, sync, D2H, free.
// Here the layout (row/column-major, transposes) is first reduced to strides;
// the larger of A and B becomes the streamed operand so one kernel family
// serves both tall-skinny and short-wide shapes; then a selector picks the
// variant and the kernels run on the caller's stream.
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_select.h"

namespace hml {

namespace {

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Strides of op(X)(i,j) for a matrix stored in `order` with leading
// dimension ld: returns (stride of i, stride of j).
void op_strides(hmlOrder order, hmlOp op, int64_t ld, int64_t *si,
                int64_t *sj) {
  const bool rowmajor_view = (order == HML_ROW_MAJOR) == (op == HML_OP_N);
  if (rowmajor_view) { *si = ld; *sj = 1; } else { *si = 1; *sj = ld; }
}

hmlStatus check_gemm_args(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M,
                          int64_t N, int64_t K, const float *A, int64_t lda,
                          const float *B, int64_t ldb, const float *C,
                          int64_t ldc) {
  if (order != HML_ROW_MAJOR && order != HML_COL_MAJOR)
    return fail(HML_STATUS_INVALID_VALUE, "order must be row- or column-major");
  if ((ta != HML_OP_N && ta != HML_OP_T) || (tb != HML_OP_N && tb != HML_OP_T))
    return fail(HML_STATUS_INVALID_VALUE, "trans must be N or T");
  if (M < 0 || N < 0 || K < 0)
    return fail(HML_STATUS_INVALID_VALUE, "negative dimension M=%lld N=%lld K=%lld",
                (long long)M, (long long)N, (long long)K);
  // BLAS leading-dimension rules: ld >= the extent of the contiguous index.
  const bool rm = order == HML_ROW_MAJOR;
  const int64_t a_inner = rm ? (ta == HML_OP_N ? K : M) : (ta == HML_OP_N ? M : K);
  const int64_t b_inner = rm ? (tb == HML_OP_N ? N : K) : (tb == HML_OP_N ? K : N);
  const int64_t c_inner = rm ? N : M;
  if (lda < std::max<int64_t>(1, a_inner))
    return fail(HML_STATUS_INVALID_VALUE, "lda=%lld too small", (long long)lda);
  if (ldb < std::max<int64_t>(1, b_inner))
    return fail(HML_STATUS_INVALID_VALUE, "ldb=%lld too small", (long long)ldb);
  if (ldc < std::max<int64_t>(1, c_inner))
    return fail(HML_STATUS_INVALID_VALUE, "ldc=%lld too small", (long long)ldc);
  if (M > 0 && N > 0 && !C) return fail(HML_STATUS_INVALID_VALUE, "C is NULL");
  if (M > 0 && N > 0 && K > 0 && (!A || !B))
    return fail(HML_STATUS_INVALID_VALUE, "A or B is NULL");
  return HML_STATUS_SUCCESS;
}

}  // namespace

GemmCanon canonicalize(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M,
                       int64_t N, int64_t K, const float *A, int64_t lda,
                       const float *B, int64_t ldb, float *C, int64_t ldc) {
  GemmCanon g;
  op_strides(order, ta, lda, &g.sAm, &g.sAk);
  op_strides(order, tb, ldb, &g.sBk, &g.sBn);
  if (order == HML_ROW_MAJOR) { g.sCm = ldc; g.sCn = 1; } else { g.sCm = 1; g.sCn = ldc; }
  g.A = A; g.B = B; g.C = C; g.M = M; g.N = N; g.K = K; g.swapped = false;
  // Stream the larger operand.  Swapping computes C^T = B^T A^T.
  if (N > M) {
    GemmCanon s = g;
    s.A = g.B; s.sAm = g.sBn; s.sAk = g.sBk;
    s.B = g.A; s.sBk = g.sAk; s.sBn = g.sAm;
    s.sCm = g.sCn; s.sCn = g.sCm;
    s.M = g.N; s.N = g.M;
    s.swapped = true;
    g = s;
  }
  // Degenerate extents make either stride acceptable.
  if (g.sAk == 1 || g.K == 1) g.form = kFormDot;
  else g.form = kFormAxpy;  // sAm == 1 for every BLAS layout here
  if (g.form == kFormDot && g.K == 1) g.sAk = 1;
  if (g.form == kFormAxpy && g.M == 1) g.sAm = 1;
  return g;
}

GemmKey make_key(const GemmCanon &g, float beta, uint32_t flags, hmlOrder order,
                 hmlOp ta, hmlOp tb) {
  GemmKey k;
  k.form = g.form;
  k.M = g.M; k.N = g.N; k.K = g.K;
  k.beta0 = (beta == 0.0f && !(flags & HML_FLAG_READ_C_ALWAYS)) ? 1 : 0;
  k.transA = ta; k.transB = tb; k.order = order;
  k.gemv_ok = g.N == 1 && (g.sBk == 1 || g.K == 1) && (g.sCm == 1 || g.M == 1);
  k.cM = g.swapped ? g.N : g.M;
  k.cN = g.swapped ? g.M : g.N;
  return k;
}

// Canonical form of a problem stored densely (leading dimension = extent of
// the contiguous index), for queries that get no pointers.
GemmCanon dense_canon(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M,
                      int64_t N, int64_t K) {
  const bool rm = order == HML_ROW_MAJOR;
  const int64_t lda = std::max<int64_t>(1, rm ? (ta == HML_OP_N ? K : M) : (ta == HML_OP_N ? M : K));
  const int64_t ldb = std::max<int64_t>(1, rm ? (tb == HML_OP_N ? N : K) : (tb == HML_OP_N ? K : N));
  const int64_t ldc = std::max<int64_t>(1, rm ? N : M);
  return canonicalize(order, ta, tb, M, N, K, nullptr, lda, nullptr, ldb,
                      nullptr, ldc);
}

namespace {

// Resolved launch geometry of a skinny variant.
//  Both forms: K split into chunks of kchunk (a multiple of the form's K
//    step: 64 for DOT, 16 for AXPY), kw consecutive chunks per workgroup
//    summed in LDS, and `ksplit` workgroups along K whose partials are
//    reduced by a second kernel when > 1; B is read in place.  variant.wpb
//    bounds kw (DOTB: waves take rows instead, kw = 1).  AXPY grids cover
//    nct_pad column tiles (a multiple of the variant's CT).
struct SkinnyPlan {
  int nct_pad;
  int64_t kchunk;
  int ksplit;   // partial sums to reduce (1: none)
  int kw;       // waves per workgroup along K
  size_t p_floats;
};

SkinnyPlan plan_skinny(const GemmKey &k, const hmlSgemmVariant &v) {
  SkinnyPlan pl;
  const int64_t nct = ceil_div(std::max<int64_t>(k.N, 1), 16);
  const int64_t Kc = std::max<int64_t>(k.K, 1);
  int64_t ks = std::max<int64_t>(1, v.ksplit);
  const int64_t step = k.form == kFormDot ? kDotStageK : kAxpyKStep;
  const int64_t kchunk = ceil_div(ceil_div(Kc, ks), step) * step;
  const int64_t nchunks = ceil_div(Kc, kchunk);
  pl.kchunk = kchunk;
  pl.nct_pad = k.form == kFormDot ? (int)nct : (int)(ceil_div(nct, v.ct) * v.ct);
  if (v.family == HML_SGEMM_FAMILY_DOTB) {
    // the waves of a workgroup take rows: every chunk is a workgroup
    pl.kw = 1;
    pl.ksplit = (int)nchunks;
  } else {
    pl.kw = (int)std::max<int64_t>(1, std::min<int64_t>(std::max(1, v.wpb), nchunks));
    if (k.form == kFormAxpy)  // power-of-two workgroups
      while (pl.kw & (pl.kw - 1)) --pl.kw;
    pl.ksplit = (int)ceil_div(nchunks, pl.kw);
  }
  pl.p_floats = pl.ksplit > 1 ? (size_t)pl.ksplit * k.M * k.N : 0;
  return pl;
}

// Every in-range byte offset a skinny wave forms from its buffer
// descriptors stays below 2^31 (the kernels' addressing contract).  The
// AXPY kernel rebuilds its descriptors per 16-k step, so only one step's
// extent counts there.
bool skinny_offsets_fit(const GemmCanon &g, const hmlSgemmVariant &v,
                        const SkinnyPlan &pl) {
  const double lim = 2147483647.0;
  if (g.form == kFormAxpy) {
    const double a = (64.0 * v.rt + 16.0 * (double)g.sAk) * 4.0;
    const double b = (16.0 * v.ct * (double)g.sBn + 16.0 * (double)g.sBk) * 4.0;
    return a < lim && b < lim;
  }
  const double rows = (double)std::min<int64_t>(16 * v.rt, g.M);
  const double cols = (double)std::min<int64_t>(16 * v.ct, g.N);
  const double a = ((rows - 1) * (double)g.sAm + pl.kchunk + 4) * 4.0;
  const double b = ((pl.kchunk + 16.0) * (double)g.sBk + (cols - 1) * (double)g.sBn) * 4.0;
  return a < lim && b < lim;
}

GemvParams gemv_params_from_canon(const GemmCanon &g, int family, float alpha,
                                  float beta, int read_c) {
  GemvParams p{};
  if (family == HML_SGEMV_FAMILY_DOT) {
    p.A = g.A; p.lda = g.sAm; p.M = g.M; p.N = g.K;
  } else {
    p.A = g.A; p.lda = g.sAk; p.M = g.K; p.N = g.M;
  }
  p.x = g.B; p.y = g.C;
  p.alpha = alpha; p.beta = beta; p.read_c = read_c;
  return p;
}

}  // namespace

// ---- variant choice -------------------------------------------------------

void heuristic_sgemm(const GemmKey &k, hmlSgemmVariant *v) {
  if (k.gemv_ok) {
    GemvKey gk{k.form == kFormDot ? 1 : 2, k.M, k.K, k.beta0};
    hmlSgemvVariant gv;
    heuristic_sgemv(gk, &gv);
    v->family = HML_SGEMM_FAMILY_GEMV;
    v->rt = gv.lpr; v->ct = gv.rb; v->ksplit = gv.split; v->wpb = gv.wpb;
    return;
  }
  if (k.N > 128 && k.M >= 256 && k.K >= 64 && (k.N > 512 || k.K <= 512) &&
      ceil_div(k.M, 128) * ceil_div(k.N, 128) >= 256) {
    // Not skinny: plain FP32 dense tiles (A re-read N/128 times otherwise),
    // when its 128 x 128 output tiles alone fill the CUs (it has no K
    // split).  Measured on MI355X: up to N = 512 the K-split skinny kernels
    // are faster once K exceeds 512 (up to 2x), and dense wins only barely
    // below that.
    *v = hmlSgemmVariant{HML_SGEMM_FAMILY_DENSE, 1, 1, 1, 4};
    return;
  }
  if (k.form == kFormDot && k.N == 2) {
    // Two columns: the multi-vector GEMV kernel with the SGEMV heuristic's
    // geometry for the same rows (measured on 8192 x 2 x 8192: 0.045 ms
    // against 0.050 for the best matrix-core tile, 0.042 against 0.049 with
    // K-contiguous B; at 4 columns the two are even, at 8 the tile wins).
    GemvKey gk{HML_SGEMV_FAMILY_DOT, k.M, k.K, k.beta0};
    hmlSgemvVariant gv;
    heuristic_sgemv(gk, &gv);
    const hmlSgemmVariant mv{HML_SGEMM_FAMILY_GEMVN, gv.lpr, std::min(gv.rb, 4), gv.split, gv.wpb};
    if (sgemm_variant_runnable(k, mv)) {
      *v = mv;
      return;
    }
  }
  const int64_t nct = ceil_div(k.N, 16);
  int64_t tiles;
  if (k.form == kFormDot) {
    // k_skinny_dot: one tile is 16 rt rows x 16 ct columns; the 4 waves of
    // a workgroup split K and sum in LDS.
    v->family = HML_SGEMM_FAMILY_DOT;
    v->ct = nct <= 1 ? 1 : nct <= 2 ? 2 : 4;
    // one or two columns: more row tiles per wave spread B's load
    // instructions over more of A (measured 6% on 8192 x 2 x 8192)
    v->rt = k.M >= 4096 && k.N <= 2 ? 4 : k.M >= 2048 ? 2 : 1;
    v->wpb = 4;
    tiles = ceil_div(k.M, 16 * v->rt) * ceil_div(nct, v->ct);
  } else {
    // k_skinny_axpy: 64 rows x 16 ct columns per wave; the 4 waves of a
    // workgroup split K and sum in LDS.  8 column tiles hold one wave per
    // SIMD: two blocks of 4 measured faster for N = 128.
    v->family = HML_SGEMM_FAMILY_AXPY;
    v->ct = nct <= 1 ? 1 : nct <= 2 ? 2 : 4;
    v->rt = 1;
    v->wpb = 4;
    tiles = ceil_div(k.M, 64) * ceil_div(nct, v->ct);
  }
  // About one resident round of waves: for the DOT kernel the number its
  // register and LDS budget lets the device hold at once (from the
  // occupancy calculator), so the whole split runs concurrently; each K
  // chunk at least 256 long.
  // AXPY form: 1024 waves (measured on 8192 x N x 8192, N = 2..128: within
  // 3% of the best variant, where 2048 waves were 5-20% slower)
  int64_t target = 1024;
  if (k.form == kFormDot) {
    const int rw = skinny_dot_resident_waves(v->rt, v->ct);
    target = rw > 0 ? rw : 3072;
  }
  int64_t ks = tiles >= target ? 1 : (target + tiles / 2) / tiles;
  ks = std::min<int64_t>(ks, std::max<int64_t>(1, k.K / 256));
  ks = std::min<int64_t>(ks, 64);
  // FP32-bound DOT tiles (N > 48) with enough tiles for a third of a
  // resident round: keep the split inside the workgroup (no partials, no
  // reduction pass) -- measured faster on 8192 x 64 x 8192 at half the
  // resident waves
  if (k.form == kFormDot && k.N > 48 && tiles * v->wpb * 3 >= target)
    ks = std::min<int64_t>(ks, v->wpb);
  v->ksplit = (int)std::max<int64_t>(1, ks);
}

void basic_sgemm(const GemmKey &k, hmlSgemmVariant *v) {
  // Fixed fallback, the counterpart of the reference's varK<6,6>
  // (src/hml_sgemm_kernel.cu:172-184): one geometry for every shape.
  v->family = k.form == kFormDot ? HML_SGEMM_FAMILY_DOT : HML_SGEMM_FAMILY_AXPY;
  v->rt = 1;
  v->ct = 4;
  v->ksplit = 1;
  v->wpb = 4;
}

bool sgemm_variant_runnable(const GemmKey &k, const hmlSgemmVariant &v) {
  if (v.family == HML_SGEMM_FAMILY_DOTB)
    return k.form == kFormDot && v.ksplit >= 1 && v.ksplit <= 4096 && v.wpb == 4 &&
           skinny_dotb_runnable(v.rt, v.ct);
  if (v.family == HML_SGEMM_FAMILY_DOT)
    return k.form == kFormDot && v.ksplit >= 1 && v.ksplit <= 4096 &&
           (v.wpb == 1 || v.wpb == 2 || v.wpb == 4 || v.wpb == 8) &&
           skinny_dot_runnable(v.rt, v.ct, v.wpb);
  if (v.family == HML_SGEMM_FAMILY_AXPY)
    return k.form == kFormAxpy && v.ksplit >= 1 && v.ksplit <= 4096 &&
           skinny_axpy_runnable(v.rt, v.ct, v.wpb);
  if (v.family == HML_SGEMM_FAMILY_GEMV) {
    if (!k.gemv_ok) return false;
    GemvKey gk{k.form == kFormDot ? 1 : 2, k.M, k.K, k.beta0};
    hmlSgemvVariant gv{gk.family, v.rt, v.ct, v.ksplit, v.wpb};
    return sgemv_variant_runnable(gk, gv);
  }
  if (v.family == HML_SGEMM_FAMILY_GEMVN)
    return k.form == kFormDot && (k.N == 2 || k.N == 4 || k.N == 8) && v.ksplit >= 1 &&
           v.ksplit <= 65535 && gemv_wpb_ok(HML_SGEMV_FAMILY_DOT, v.wpb) &&
           gemv_multi_has_variant(v.rt, v.ct, (int)k.N);
  if (v.wpb != 1 && v.wpb != 2 && v.wpb != 4) return false;
  if (is_dense_family(v.family)) return true;
  return false;
}

void select_sgemm(const GemmKey &k, hmlSgemmVariant *v, int *source) {
  // A table rule for this exact shape bucket is the exhaustive autotuner's
  // own measurement: it outranks the learned selector, which generalises
  // from other shapes; a more distant rule ranks below it.
  if (config_select_sgemm(k, v, source, true) && sgemm_variant_runnable(k, *v)) return;
  if (learned_select_sgemm(k, v) && sgemm_variant_runnable(k, *v)) {
    *source = 2;
    return;
  }
  if (config_select_sgemm(k, v, source) && sgemm_variant_runnable(k, *v)) return;
  heuristic_sgemm(k, v);
  *source = 0;
}

// ---- execution ---------------------------------------------------------------

namespace {

size_t workspace_bytes(const GemmKey &k, const hmlSgemmVariant &v) {
  if (is_dense_family(v.family))
    return dense_workspace(dense_precision(v.family), k.M, k.N, k.K);
  if (v.family == HML_SGEMM_FAMILY_GEMV || v.family == HML_SGEMM_FAMILY_GEMVN) {
    GemvParams p{};
    const int slots = gemv_resolve_split(k.form == kFormDot ? HML_SGEMV_FAMILY_DOT : HML_SGEMV_FAMILY_AXPY,
                                         k.K, v.rt, std::max(1, v.ksplit), v.wpb, &p);
    const int64_t cols = v.family == HML_SGEMM_FAMILY_GEMVN ? k.N : 1;
    return slots > 1 ? align256((size_t)slots * k.M * cols * sizeof(float)) : 0;
  }
  const SkinnyPlan pl = plan_skinny(k, v);
  return align256(pl.p_floats * sizeof(float));
}

hmlStatus run_gemv_family(const GemmCanon &g, const GemmKey &k,
                          const hmlSgemmVariant &v, float alpha, float beta,
                          int read_c, void *ws, hipStream_t st);

// Strides too large for the 31-bit buffer offsets of the skinny and
// multi-vector kernels: plain FP32 dense tiles instead (they address with
// 64-bit pointers through the split pass).  A caller's workspace is used
// when it holds the dense kernel's operand copies; without one, the
// library workspace of the stream.
hmlStatus run_dense_fallback(const GemmCanon &g, const GemmKey &k, float alpha, float beta,
                             int read_c, void *caller_ws, size_t ws_bytes, hipStream_t st) {
  const size_t dn = dense_workspace(0, k.M, k.N, k.K);
  WsLease lease;
  void *dws = caller_ws;
  if (caller_ws) {
    if (ws_bytes < dn)
      return fail(HML_STATUS_WORKSPACE_TOO_SMALL,
                  "strides exceed the skinny kernels' 31-bit offsets; the dense fallback "
                  "needs a workspace of %zu bytes (got %zu)", dn, ws_bytes);
  } else {
    hmlStatus s = library_workspace(dn, st, &lease);
    if (s != HML_STATUS_SUCCESS) return s;
    dws = lease.ptr();
  }
  return run_dense(0, g.A, g.sAm, g.sAk, g.B, g.sBk, g.sBn, g.C, g.sCm, g.sCn, g.M, g.N, g.K,
                   alpha, beta, read_c, dws, dn, st);
}

hmlStatus run_sgemm(const GemmCanon &g, const GemmKey &k,
                    const hmlSgemmVariant &v, float alpha, float beta,
                    int read_c, void *ws, size_t ws_bytes, hipStream_t st) {
  const size_t need = workspace_bytes(k, v);
  void *const caller_ws = ws;
  WsLease lease;  // held until the kernels below are queued
  if (need > 0 && !ws) {
    hmlStatus s = library_workspace(need, st, &lease);
    if (s != HML_STATUS_SUCCESS) return s;
    ws = lease.ptr();
    ws_bytes = need;
  } else if (need > ws_bytes) {
    return fail(HML_STATUS_WORKSPACE_TOO_SMALL, "workspace %zu < %zu bytes",
                ws_bytes, need);
  }
  if (v.family == HML_SGEMM_FAMILY_GEMV)
    return run_gemv_family(g, k, v, alpha, beta, read_c, ws, st);
  if (v.family == HML_SGEMM_FAMILY_GEMVN) {
    GemvNParams p{};
    p.A = g.A; p.lda = g.sAm;
    p.B = g.B; p.sBk = g.sBk; p.sBn = g.sBn;
    p.C = g.C; p.sCm = g.sCm; p.sCn = g.sCn;
    p.P = (float *)ws;
    p.M = g.M; p.K = g.K;
    p.alpha = alpha; p.beta = beta; p.read_c = read_c;
    GemvParams sp{};
    const int slots = gemv_resolve_split(HML_SGEMV_FAMILY_DOT, g.K, v.rt, std::max(1, v.ksplit),
                                         v.wpb, &sp);
    p.chunk = sp.chunk; p.split = sp.split; p.wsum = sp.wsum;
    const bool bvec = g.sBk == 1 || g.K == 1;
    if (!gemv_multi_offsets_fit(v.rt, v.ct, (int)g.N, bvec, p))
      return run_dense_fallback(g, k, alpha, beta, read_c, caller_ws, ws_bytes, st);
    hipError_t e = launch_gemv_multi(v.rt, v.ct, (int)g.N, bvec, v.wpb, p, st);
    if (e != hipSuccess) return from_hip(e, "multi-vector gemv launch");
    if (slots > 1) {
      e = launch_reduce(p.P, slots, g.M * g.N, g.N, g.C, g.sCm, g.sCn, alpha, beta, read_c, st);
      if (e != hipSuccess) return from_hip(e, "split-K reduce");
    }
    return HML_STATUS_SUCCESS;
  }
  if (is_dense_family(v.family))
    return run_dense(dense_precision(v.family), g.A, g.sAm, g.sAk, g.B, g.sBk,
                     g.sBn, g.C, g.sCm, g.sCn, g.M, g.N, g.K, alpha, beta,
                     read_c, ws, ws ? need : 0, st);

  const SkinnyPlan pl = plan_skinny(k, v);
  if (!skinny_offsets_fit(g, v, pl))
    return run_dense_fallback(g, k, alpha, beta, read_c, caller_ws, ws_bytes, st);
  hipError_t e;
  if (k.form == kFormDot) {
    DotParams p{};
    p.A = g.A; p.sAm = g.sAm;
    p.B = g.B; p.sBk = g.sBk; p.sBn = g.sBn;
    p.C = g.C; p.sCm = g.sCm; p.sCn = g.sCn;
    p.P = (float *)ws;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.kchunk = pl.kchunk;
    p.kgroups = pl.ksplit;
    p.alpha = alpha; p.beta = beta; p.read_c = read_c;
    e = v.family == HML_SGEMM_FAMILY_DOTB
            ? launch_skinny_dotb(v.rt, v.ct, g.sBk == 1, p, st)
            : launch_skinny_dot(v.rt, v.ct, g.sBk == 1, pl.kw, p, st);
  } else {
    SkinnyParams p{};
    p.A = g.A; p.sAm = g.sAm; p.sAk = g.sAk;
    p.B = g.B; p.sBk = g.sBk; p.sBn = g.sBn;
    p.C = g.C; p.sCm = g.sCm; p.sCn = g.sCn;
    p.P = (float *)ws;
    p.M = g.M; p.N = g.N; p.K = g.K;
    p.kchunk = pl.kchunk; p.kgroups = pl.ksplit;
    p.nct_pad = pl.nct_pad;
    p.alpha = alpha; p.beta = beta; p.read_c = read_c;
    e = launch_skinny_axpy(v.rt, v.ct, g.M % 4 == 0, g.sBk == 1 || g.K == 1, pl.kw, p, st);
  }
  if (e != hipSuccess) return from_hip(e, "skinny kernel launch");
  if (pl.ksplit > 1) {
    e = launch_reduce((const float *)ws, pl.ksplit, g.M * g.N, g.N, g.C, g.sCm, g.sCn,
                      alpha, beta, read_c, st);
    if (e != hipSuccess) return from_hip(e, "split-K reduce");
  }
  return HML_STATUS_SUCCESS;
}

hmlStatus run_gemv_family(const GemmCanon &g, const GemmKey &k,
                          const hmlSgemmVariant &v, float alpha, float beta,
                          int read_c, void *ws, hipStream_t st) {
  const int family = k.form == kFormDot ? HML_SGEMV_FAMILY_DOT : HML_SGEMV_FAMILY_AXPY;
  GemvParams p = gemv_params_from_canon(g, family, alpha, beta, read_c);
  p.P = (float *)ws;
  const int64_t reduced = family == HML_SGEMV_FAMILY_DOT ? p.N : p.M;
  const int64_t outputs = family == HML_SGEMV_FAMILY_DOT ? p.M : p.N;
  const int slots = gemv_resolve_split(family, reduced, v.rt, std::max(1, v.ksplit), v.wpb, &p);
  bool vec = ((uintptr_t)p.A % 16 == 0) && (p.lda % 4 == 0) && (p.N % 4 == 0) &&
             (family != HML_SGEMV_FAMILY_DOT || (uintptr_t)p.x % 16 == 0);
  hipError_t e = launch_gemv(family, v.rt, v.ct, vec, v.wpb, p, st);
  if (e != hipSuccess) return from_hip(e, "gemv kernel launch");
  if (slots > 1) {
    e = launch_reduce(p.P, slots, outputs, 1, p.y, 1, 0, alpha, beta,
                      read_c, st);
    if (e != hipSuccess) return from_hip(e, "gemv split reduce");
  }
  return HML_STATUS_SUCCESS;
}

hmlStatus sgemm_impl(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M, int64_t N,
                     int64_t K, float alpha, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float beta, float *C,
                     int64_t ldc, uint32_t flags, const hmlSgemmVariant *forced,
                     void *ws, size_t ws_bytes, hipStream_t st,
                     hmlSgemmVariant *used) {
  clear_error();
  hmlStatus s = check_gemm_args(order, ta, tb, M, N, K, A, lda, B, ldb, C, ldc);
  if (s != HML_STATUS_SUCCESS) return s;
  if (M == 0 || N == 0) return HML_STATUS_SUCCESS;
  const int read_c = (beta != 0.0f || (flags & HML_FLAG_READ_C_ALWAYS)) ? 1 : 0;
  GemmCanon g = canonicalize(order, ta, tb, M, N, K, A, lda, B, ldb, C, ldc);
  if (K == 0 || (alpha == 0.0f && !(flags & HML_FLAG_READ_C_ALWAYS))) {
    // No product term: C = alpha*0 + beta*C (BLAS quick return; with the
    // reference semantics alpha == 0 still runs the product so that
    // 0 * Inf gives NaN exactly as the reference kernels do).
    if (read_c && beta == 1.0f) return HML_STATUS_SUCCESS;
    return from_hip(launch_scale(g.C, g.M, g.N, g.sCm, g.sCn, alpha, beta,
                                 read_c, st),
                    "scale C");
  }
  GemmKey key = make_key(g, beta, flags, order, ta, tb);
  hmlSgemmVariant v;
  if (forced) {
    v = *forced;
    if (!sgemm_variant_runnable(key, v))
      return fail(HML_STATUS_NO_KERNEL,
                  "variant family=%d rt=%d ct=%d ksplit=%d wpb=%d cannot run form %d",
                  v.family, v.rt, v.ct, v.ksplit, v.wpb, key.form);
  } else {
    int src;
    select_sgemm(key, &v, &src);
  }
  if (used) *used = v;
  return run_sgemm(g, key, v, alpha, beta, read_c, ws, ws_bytes, st);
}

// Host-pointer compatibility path (src/hml_sgemm.cu:39-113,120-211).
void sgemm_compat(float *C, const float *A, const float *B, int M, int N,
                  int K, float alpha, float beta, bool basic) {
  if (M < 0 || N < 0 || K < 0) {
    fprintf(stderr, "; Error: invalid SGEMM dimensions\n");
    exit(1);
  }
  const size_t szA = (size_t)M * K, szB = (size_t)K * N, szC = (size_t)M * N;
  float *dA = nullptr, *dB = nullptr, *dC = nullptr;
  std::lock_guard<std::mutex> lk(compat_mutex());
  hmlStatus s = compat_buffers(szA * 4, szB * 4, szC * 4, &dA, &dB, &dC);
  if (s != HML_STATUS_SUCCESS) compat_die("alloc A, B, C", s);
  if ((s = compat_h2d(dA, A, szA * 4)) != HML_STATUS_SUCCESS) compat_die("copy A", s);
  if ((s = compat_h2d(dB, B, szB * 4)) != HML_STATUS_SUCCESS) compat_die("copy B", s);
  if ((s = compat_h2d(dC, C, szC * 4)) != HML_STATUS_SUCCESS) compat_die("copy C", s);
  hmlSgemmVariant v;
  GemmCanon g = canonicalize(HML_ROW_MAJOR, HML_OP_N, HML_OP_N, M, N, K, dA,
                             std::max(K, 1), dB, std::max(N, 1), dC, std::max(N, 1));
  GemmKey key = make_key(g, beta, HML_FLAG_READ_C_ALWAYS, HML_ROW_MAJOR,
                         HML_OP_N, HML_OP_N);
  int src = 0;
  if (basic) basic_sgemm(key, &v); else select_sgemm(key, &v, &src);
  if (verbosity() >= 1 && !basic)
    fprintf(stderr, "Info: kernel type = %d, colStops = %d, rowStops = %d\n",
            v.family, v.ct, v.rt);
  s = sgemm_impl(HML_ROW_MAJOR, HML_OP_N, HML_OP_N, M, N, K, alpha, dA,
                 std::max(K, 1), dB, std::max(N, 1), beta, dC, std::max(N, 1),
                 HML_FLAG_READ_C_ALWAYS, &v, nullptr, 0, nullptr, nullptr);
  if (s != HML_STATUS_SUCCESS) {
    fprintf(stderr, "Run kernel: %s\n", hmlLastError());
    exit(1);
  }
  if ((s = hmlDeviceSynchronize()) != HML_STATUS_SUCCESS) {
    fprintf(stderr, "Run kernel: %s\n", hmlLastError());
    exit(1);
  }
  if ((s = compat_d2h(C, dC, szC * 4)) != HML_STATUS_SUCCESS) {
    fprintf(stderr, "Copy C off of device: %s\n", hmlLastError());
    exit(1);
  }
}

}  // namespace
}  // namespace hml

using namespace hml;

extern "C" {

hmlStatus hmlSgemmEx(hmlOrder order, hmlOp transA, hmlOp transB, int64_t M,
                     int64_t N, int64_t K, float alpha, const float *A,
                     int64_t lda, const float *B, int64_t ldb, float beta,
                     float *C, int64_t ldc, uint32_t flags, void *workspace,
                     size_t ws_bytes, void *stream) {
  return sgemm_impl(order, transA, transB, M, N, K, alpha, A, lda, B, ldb,
                    beta, C, ldc, flags, nullptr, workspace, ws_bytes,
                    (hipStream_t)stream, nullptr);
}

hmlStatus hmlSgemmExVariant(hmlOrder order, hmlOp transA, hmlOp transB,
                            int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, float beta, float *C, int64_t ldc,
                            uint32_t flags, const hmlSgemmVariant *variant,
                            void *workspace, size_t ws_bytes, void *stream) {
  if (!variant) return fail(HML_STATUS_INVALID_VALUE, "variant is NULL");
  return sgemm_impl(order, transA, transB, M, N, K, alpha, A, lda, B, ldb,
                    beta, C, ldc, flags, variant, workspace, ws_bytes,
                    (hipStream_t)stream, nullptr);
}

size_t hmlSgemmWorkspaceSize(hmlOrder order, hmlOp transA, hmlOp transB,
                             int64_t M, int64_t N, int64_t K,
                             const hmlSgemmVariant *variant) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  GemmCanon g = dense_canon(order, transA, transB, M, N, K);
  // The strides and beta are unknown here: when N == 1 after orientation
  // the call may take either the GEMV or the skinny route, and a selector
  // may choose differently for beta == 0 and beta != 0.  Return the largest
  // need over those cases.
  size_t need = 0;
  for (int beta_nz = 0; beta_nz <= 1; ++beta_nz) {
    GemmKey key = make_key(g, beta_nz ? 1.0f : 0.0f, 0, order, transA, transB);
    for (int gemv = 0; gemv <= (g.N == 1 ? 1 : 0); ++gemv) {
      key.gemv_ok = gemv == 1;
      hmlSgemmVariant v;
      if (variant) {
        v = *variant;
      } else {
        int src;
        select_sgemm(key, &v, &src);
      }
      if (!variant || sgemm_variant_runnable(key, v))
        need = std::max(need, workspace_bytes(key, v));
    }
  }
  return need;
}

hmlStatus hmlSgemmSelect(hmlOrder order, hmlOp transA, hmlOp transB,
                         int64_t M, int64_t N, int64_t K, float beta,
                         hmlSgemmVariant *variant, int32_t *source) {
  if (!variant) return fail(HML_STATUS_INVALID_VALUE, "variant is NULL");
  if (M <= 0 || N <= 0 || K <= 0)
    return fail(HML_STATUS_INVALID_VALUE, "dimensions must be positive");
  GemmCanon g = dense_canon(order, transA, transB, M, N, K);
  GemmKey key = make_key(g, beta, 0, order, transA, transB);
  int src = 0;
  select_sgemm(key, variant, &src);
  if (source) *source = src;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemmCheckF64(hmlOrder order, hmlOp transA, hmlOp transB,
                           int64_t M, int64_t N, int64_t K, float alpha,
                           const float *A, int64_t lda, const float *B,
                           int64_t ldb, float beta, const float *C0,
                           const float *C, int64_t ldc, double *stats) {
  clear_error();
  hmlStatus s = check_gemm_args(order, transA, transB, M, N, K, A, lda, B, ldb, C, ldc);
  if (s != HML_STATUS_SUCCESS) return s;
  if (!stats) return fail(HML_STATUS_INVALID_VALUE, "stats is NULL");
  for (int i = 0; i < 6; ++i) stats[i] = 0.0;
  if (M == 0 || N == 0) return HML_STATUS_SUCCESS;
  CheckParams p{};
  op_strides(order, transA, lda, &p.sAm, &p.sAk);
  op_strides(order, transB, ldb, &p.sBk, &p.sBn);
  if (order == HML_ROW_MAJOR) { p.sCm = ldc; p.sCn = 1; } else { p.sCm = 1; p.sCn = ldc; }
  p.A = A; p.B = B; p.C0 = beta != 0.0f ? C0 : nullptr; p.C = C;
  p.M = M; p.N = N; p.K = K; p.alpha = alpha; p.beta = beta;
  if (beta != 0.0f && !C0) return fail(HML_STATUS_INVALID_VALUE, "C0 needed when beta != 0");
  const int64_t nblk = ((M + 15) / 16) * ((N + 15) / 16);
  double *dstats = nullptr;
  hipError_t e = hipMalloc(&dstats, (size_t)nblk * 6 * sizeof(double));
  if (e != hipSuccess) return from_hip(e, "hipMalloc(check stats)");
  p.stats = dstats;
  int64_t nb = 0;
  e = launch_check_f64(p, &nb, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(dstats);
    return from_hip(e, "check kernel");
  }
  double *h = (double *)malloc((size_t)nblk * 6 * sizeof(double));
  if (!h) { (void)hipFree(dstats); return fail(HML_STATUS_ALLOC_FAILED, "host stats"); }
  e = hipMemcpy(h, dstats, (size_t)nblk * 6 * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(dstats);
  if (e != hipSuccess) { free(h); return from_hip(e, "copy check stats"); }
  for (int64_t b = 0; b < nblk; ++b) {
    stats[0] += h[b * 6 + 0];
    stats[1] += h[b * 6 + 1];
    stats[2] += h[b * 6 + 2];
    stats[3] = std::max(stats[3], h[b * 6 + 3]);
    stats[4] = std::max(stats[4], h[b * 6 + 4]);
    stats[5] += h[b * 6 + 5];
  }
  free(h);
  return HML_STATUS_SUCCESS;
}

void hmlSgemmBasic(float *C, const float *A, const float *B, const int M,
                   const int N, const int K, const float alpha,
                   const float beta) {
  sgemm_compat(C, A, B, M, N, K, alpha, beta, true);
}

void hmlSgemm(float *C, const float *A, const float *B, const int M,
              const int N, const int K, const float alpha, const float beta) {
  sgemm_compat(C, A, B, M, N, K, alpha, beta, false);
}

}  // extern "C"
