// hml_gemv.cpp -- SGEMV dispatch: reference-compatible hmlSgemv /
// hmlSgemvBasic (host pointers, exit on error) and the extended
// device-pointer entry points.
//
// Reference flow being replaced: src/hml_sgemv.cu:125-213 (lazy init, lazy
// online autotune, alloc + copy + bind, hmlSgemvKernelSelect
// src/hml_sgemv_kernel.cu:315-355, launch, sync, copy back).  Every layout is
// reduced to one of two forms: DOT (each output is a contiguous row of A) or
// AXPY (each output is a column, rows of A contiguous).
#include <stdio.h>

#include <algorithm>

#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_select.h"

namespace hml {

namespace {

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int lpr_for(int64_t len) {
  // Lanes per row / column chunk: enough lanes that one 128-bit load per
  // lane covers the row (or column chunk), between 4 and 64.
  int lpr = 4;
  while (lpr < 64 && (int64_t)lpr * 4 < len) lpr <<= 1;
  return lpr;
}

struct GemvCanon {
  GemvParams p;   // without split / chunk / partials
  int family;     // 1 DOT, 2 AXPY
  int64_t outputs, reduced;
};

GemvCanon canonicalize(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                       const float *A, int64_t lda, const float *x, float *y) {
  GemvCanon c;
  GemvParams &p = c.p;
  p = GemvParams{};
  p.A = A; p.lda = lda; p.x = x; p.y = y;
  // Row-major N and column-major T: outputs are contiguous rows of A.
  const bool dot = (order == HML_ROW_MAJOR) == (trans == HML_OP_N);
  const int64_t out = trans == HML_OP_N ? M : N;
  const int64_t red = trans == HML_OP_N ? N : M;
  if (dot) {
    c.family = HML_SGEMV_FAMILY_DOT;
    p.M = out; p.N = red;
  } else {
    c.family = HML_SGEMV_FAMILY_AXPY;
    p.M = red; p.N = out;
  }
  c.outputs = out; c.reduced = red;
  return c;
}

}  // namespace

GemvKey gemv_key(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                 int beta0) {
  GemvCanon c = canonicalize(order, trans, M, N, nullptr, 1, nullptr, nullptr);
  return GemvKey{c.family, c.outputs, c.reduced, beta0};
}

int ilog2_floor(int64_t x) {
  int l = 0;
  while (x > 1) { x >>= 1; ++l; }
  return l;
}

void heuristic_sgemv(const GemvKey &k, hmlSgemvVariant *v) {
  // Fitted to steady-state sweeps of every candidate on MI355X (16384^2,
  // 8192^2, 65536 x 1023, 2^20 x 64, both forms): 128-bit loads, thin waves
  // and many of them in flight.
  v->family = k.family;
  if (k.family == HML_SGEMV_FAMILY_DOT) {
    // Lanes per row: one 128-bit load per lane up to 16 floats, two at 64,
    // four from 128; rows of 512 .. 2047 floats on 32 lanes, longer on 64.
    const int64_t loads = std::min<int64_t>(4, std::max<int64_t>(1, k.reduced / 32));
    v->lpr = lpr_for(ceil_div(k.reduced, loads));
    if (k.reduced >= 512 && k.reduced < 2048) v->lpr = 32;
    v->wpb = 2;
    v->rb = v->lpr >= 32 ? (k.outputs >= 8192 ? 2 : 1)
                         : (v->lpr <= 8 && k.outputs >= ((int64_t)1 << 17) ? 4 : 1);
    // Split the rows into enough pieces for about 64K waves when A is at
    // least 512 MB, 4K below that (gemv_split_rule).  The two waves of a
    // workgroup add their chunks in LDS.  Measured across 16384^2, 32768^2,
    // 65536 x 8192 and 8192 x 65536: this two-row, two-wave geometry is the
    // best or within 3% of it; workgroups of 8 waves covering the whole
    // split were up to 7% slower, single-row waves on a 256K-wave split 3%
    // faster on 16384^2 only and up to 7% slower on the others.
    v->split = gemv_split_rule(k, v->lpr, v->rb);
  } else {
    // Column chunks of at most 128 floats (32 lanes), 8-wave workgroups
    // (their waves' sums meet in LDS); the rows split so that about 256
    // workgroups (2048 waves) run -- 1024 below 64 MB of A (gemv_split_rule).
    // Measured against 4-wave workgroups on twice the splits over seven
    // shapes from 4099 x 2049 to 32768^2: 6% faster to 3% slower (16384^2:
    // 3% faster).
    v->lpr = std::min(32, lpr_for(k.outputs));
    v->rb = 1;
    v->wpb = 8;
    v->split = gemv_split_rule(k, v->lpr, v->rb);
  }
}

int gemv_split_rule(const GemvKey &k, int lpr, int rb) {
  if (k.family == HML_SGEMV_FAMILY_DOT) {
    // DOT: about 64K waves when A is at least 512 MB, 4K below that (a
    // split adds partials and a reduction pass, which costs more than it
    // hides on mid-size problems), each piece at least 4 loads per lane.
    const int64_t rows_per_wave = (int64_t)(64 / lpr) * rb;
    const int64_t waves = ceil_div(k.outputs, rows_per_wave);
    const int64_t target = (double)k.outputs * (double)k.reduced * 4.0 >= 536870912.0 ? 65536 : 4096;
    int64_t split = 1;
    if (waves < target) {
      split = ceil_div(target, waves);
      split = std::min<int64_t>(split, std::max<int64_t>(1, k.reduced / (4 * lpr * 4)));
    }
    return (int)std::max<int64_t>(1, std::min<int64_t>(split, 1024));
  }
  // AXPY: about 256 workgroups -- 1024 below 64 MB of A, where each wave's
  // short row range is latency-bound (4099 x 2049: 18.5 -> 12 us) -- each
  // split at least 64 rows long, at most 512 splits.
  const int64_t ncb = ceil_div(k.outputs, 4 * lpr);
  const int64_t target_wg = (double)k.outputs * (double)k.reduced * 4.0 < 67108864.0 ? 1024 : 256;
  int64_t split = ceil_div(target_wg, ncb);
  split = std::min<int64_t>(split, std::max<int64_t>(1, k.reduced / 64));
  return (int)std::max<int64_t>(1, std::min<int64_t>(split, 512));
}

void basic_sgemv(const GemvKey &k, hmlSgemvVariant *v) {
  // Fixed fallback, the counterpart of the reference's basic[8]
  // (src/hml_sgemv_kernel.cu:294-310): whole-wave rows, no split.
  v->family = k.family;
  v->lpr = 64;
  v->rb = k.family == HML_SGEMV_FAMILY_DOT ? 1 : 1;
  v->split = k.family == HML_SGEMV_FAMILY_DOT ? 1 : (int)std::max<int64_t>(1, std::min<int64_t>(k.reduced / 1024, 1024));
  v->wpb = 4;
}

int gemv_resolve_split(int family, int64_t reduced, int lpr, int split, int wpb,
                       GemvParams *p) {
  const int64_t gran = family == HML_SGEMV_FAMILY_DOT ? 4 * lpr : 1;
  const int64_t r = std::max<int64_t>(reduced, 1);
  int64_t chunk = ceil_div(r, std::max(1, split));
  chunk = ceil_div(chunk, gran) * gran;
  p->chunk = chunk;
  p->split = (int)ceil_div(r, chunk);
  p->wsum = 0;
  if (family == HML_SGEMV_FAMILY_DOT && wpb > 1 && p->split > 1) {
    // whole workgroups per row block; trailing chunks past the row are
    // empty and add zero
    p->split = (int)(ceil_div(p->split, wpb) * wpb);
    p->wsum = 1;
    return p->split / wpb;
  }
  return p->split;
}

bool sgemv_variant_runnable(const GemvKey &k, const hmlSgemvVariant &v) {
  if (v.family != k.family) return false;
  if (!gemv_wpb_ok(v.family, v.wpb)) return false;
  if (v.split < 1 || v.split > 65535) return false;
  return gemv_has_variant(v.family, v.lpr, v.rb);
}

void select_sgemv(const GemvKey &k, hmlSgemvVariant *v, int *source) {
  // same order as select_sgemm: exact table bucket, learned selector,
  // nearest table rule, heuristic
  if (config_select_sgemv(k, v, source, true) && sgemv_variant_runnable(k, *v)) return;
  if (learned_select_sgemv(k, v) && sgemv_variant_runnable(k, *v)) {
    *source = 2;
    return;
  }
  if (config_select_sgemv(k, v, source) && sgemv_variant_runnable(k, *v)) return;
  heuristic_sgemv(k, v);
  *source = 0;
}

namespace {

size_t gemv_workspace(const GemvKey &k, const hmlSgemvVariant &v) {
  GemvParams p{};
  const int slots = gemv_resolve_split(k.family, k.reduced, v.lpr, v.split, v.wpb, &p);
  return slots > 1 ? align256((size_t)slots * k.outputs * sizeof(float)) : 0;
}

hmlStatus sgemv_impl(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                     float alpha, const float *A, int64_t lda, const float *x,
                     int64_t incx, float beta, float *y, int64_t incy,
                     uint32_t flags, const hmlSgemvVariant *forced, void *ws,
                     size_t ws_bytes, hipStream_t st) {
  clear_error();
  if (order != HML_ROW_MAJOR && order != HML_COL_MAJOR)
    return fail(HML_STATUS_INVALID_VALUE, "order must be row- or column-major");
  if (trans != HML_OP_N && trans != HML_OP_T)
    return fail(HML_STATUS_INVALID_VALUE, "trans must be N or T");
  if (M < 0 || N < 0) return fail(HML_STATUS_INVALID_VALUE, "negative dimension");
  const int64_t inner = order == HML_ROW_MAJOR ? N : M;
  if (lda < std::max<int64_t>(1, inner))
    return fail(HML_STATUS_INVALID_VALUE, "lda=%lld too small", (long long)lda);
  if (incx != 1 || incy != 1)
    return fail(HML_STATUS_UNSUPPORTED, "only incx == incy == 1 is supported");
  GemvCanon c = canonicalize(order, trans, M, N, A, lda, x, y);
  if (c.outputs == 0) return HML_STATUS_SUCCESS;
  if (!y) return fail(HML_STATUS_INVALID_VALUE, "y is NULL");
  const int read_c = (beta != 0.0f || (flags & HML_FLAG_READ_C_ALWAYS)) ? 1 : 0;
  if (c.reduced == 0 || (alpha == 0.0f && !(flags & HML_FLAG_READ_C_ALWAYS))) {
    if (read_c && beta == 1.0f) return HML_STATUS_SUCCESS;
    return from_hip(launch_scale(y, c.outputs, 1, 1, 0, alpha, beta, read_c, st),
                    "scale y");
  }
  if (!A || !x) return fail(HML_STATUS_INVALID_VALUE, "A or x is NULL");
  GemvKey key{c.family, c.outputs, c.reduced,
              (beta == 0.0f && !(flags & HML_FLAG_READ_C_ALWAYS)) ? 1 : 0};
  hmlSgemvVariant v;
  if (forced) {
    v = *forced;
    if (!sgemv_variant_runnable(key, v))
      return fail(HML_STATUS_NO_KERNEL, "gemv variant family=%d lpr=%d rb=%d split=%d wpb=%d cannot run family %d",
                  v.family, v.lpr, v.rb, v.split, v.wpb, key.family);
  } else {
    int src;
    select_sgemv(key, &v, &src);
  }
  GemvParams p = c.p;
  p.alpha = alpha; p.beta = beta; p.read_c = read_c;
  const int slots = gemv_resolve_split(c.family, c.reduced, v.lpr, v.split, v.wpb, &p);
  const size_t need = slots > 1 ? align256((size_t)slots * c.outputs * sizeof(float)) : 0;
  WsLease lease;  // held until the kernels below are queued
  if (need > 0) {
    if (!ws) {
      hmlStatus s = library_workspace(need, st, &lease);
      if (s != HML_STATUS_SUCCESS) return s;
      ws = lease.ptr();
    } else if (ws_bytes < need) {
      return fail(HML_STATUS_WORKSPACE_TOO_SMALL, "workspace %zu < %zu bytes", ws_bytes, need);
    }
  }
  p.P = (float *)ws;
  // The streaming kernels need whole 4-float groups along the contiguous
  // extent (p.N in both forms) and 16-byte aligned rows.
  const bool vec = ((uintptr_t)A % 16 == 0) && (lda % 4 == 0) && (p.N % 4 == 0) &&
                   (c.family != HML_SGEMV_FAMILY_DOT || (uintptr_t)x % 16 == 0);
  hipError_t e = launch_gemv(v.family, v.lpr, v.rb, vec, v.wpb, p, st);
  if (e != hipSuccess) return from_hip(e, "gemv kernel launch");
  if (slots > 1) {
    e = launch_reduce(p.P, slots, c.outputs, 1, y, 1, 0, alpha, beta, read_c, st);
    if (e != hipSuccess) return from_hip(e, "gemv split reduce");
  }
  return HML_STATUS_SUCCESS;
}

void sgemv_compat(float *y, const float *A, const float *x, int M, int N,
                  float alpha, float beta, bool basic) {
  if (M < 0 || N < 0) {
    fprintf(stderr, "; Error: invalid SGEMV dimensions\n");
    exit(1);
  }
  const size_t szA = (size_t)M * N;
  float *dA = nullptr, *dx = nullptr, *dy = nullptr;
  std::lock_guard<std::mutex> lk(compat_mutex());
  hmlStatus s = compat_buffers(szA * 4, (size_t)N * 4, (size_t)M * 4, &dA, &dx, &dy);
  if (s != HML_STATUS_SUCCESS) compat_die("alloc A, x, y", s);
  if ((s = compat_h2d(dA, A, szA * 4)) != HML_STATUS_SUCCESS) compat_die("copy A", s);
  if ((s = compat_h2d(dx, x, (size_t)N * 4)) != HML_STATUS_SUCCESS) compat_die("copy x", s);
  if ((s = compat_h2d(dy, y, (size_t)M * 4)) != HML_STATUS_SUCCESS) compat_die("copy y", s);
  GemvKey key{HML_SGEMV_FAMILY_DOT, M, N, 0};
  hmlSgemvVariant v;
  int src = 0;
  if (basic) basic_sgemv(key, &v); else select_sgemv(key, &v, &src);
  if (verbosity() >= 1 && !basic)
    fprintf(stderr, "Info: kernel type = %d, blockStops = %d, rowStops = %d\n",
            v.family, v.lpr / 16 > 0 ? v.lpr / 16 : 1, v.rb);
  s = sgemv_impl(HML_ROW_MAJOR, HML_OP_N, M, N, alpha, dA, std::max(N, 1), dx,
                 1, beta, dy, 1, HML_FLAG_READ_C_ALWAYS, &v, nullptr, 0, nullptr);
  if (s == HML_STATUS_SUCCESS) s = hmlDeviceSynchronize();
  if (s != HML_STATUS_SUCCESS) {
    fprintf(stderr, "Run kernel: %s\n", hmlLastError());
    exit(1);
  }
  if ((s = compat_d2h(y, dy, (size_t)M * 4)) != HML_STATUS_SUCCESS) {
    fprintf(stderr, "Copy y off of device: %s\n", hmlLastError());
    exit(1);
  }
}

}  // namespace
}  // namespace hml

using namespace hml;

extern "C" {

hmlStatus hmlSgemvEx(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                     float alpha, const float *A, int64_t lda, const float *x,
                     int64_t incx, float beta, float *y, int64_t incy,
                     uint32_t flags, void *workspace, size_t ws_bytes,
                     void *stream) {
  return sgemv_impl(order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy,
                    flags, nullptr, workspace, ws_bytes, (hipStream_t)stream);
}

hmlStatus hmlSgemvExVariant(hmlOrder order, hmlOp trans, int64_t M,
                            int64_t N, float alpha, const float *A,
                            int64_t lda, const float *x, int64_t incx,
                            float beta, float *y, int64_t incy,
                            uint32_t flags, const hmlSgemvVariant *variant,
                            void *workspace, size_t ws_bytes, void *stream) {
  if (!variant) return fail(HML_STATUS_INVALID_VALUE, "variant is NULL");
  return sgemv_impl(order, trans, M, N, alpha, A, lda, x, incx, beta, y, incy,
                    flags, variant, workspace, ws_bytes, (hipStream_t)stream);
}

size_t hmlSgemvWorkspaceSize(hmlOrder order, hmlOp trans, int64_t M,
                             int64_t N, const hmlSgemvVariant *variant) {
  if (M <= 0 || N <= 0) return 0;
  GemvCanon c = canonicalize(order, trans, M, N, nullptr, 1, nullptr, nullptr);
  // beta is unknown here and a selector may choose differently for
  // beta == 0: the larger need of the two.  The launch may use fewer splits
  // than requested, never more.
  size_t need = 0;
  for (int beta0 = 0; beta0 <= 1; ++beta0) {
    GemvKey key{c.family, c.outputs, c.reduced, beta0};
    hmlSgemvVariant v;
    if (variant) {
      v = *variant;
    } else {
      int src;
      select_sgemv(key, &v, &src);
    }
    need = std::max(need, gemv_workspace(key, v));
  }
  return need;
}

hmlStatus hmlSgemvSelect(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                         float beta, hmlSgemvVariant *variant,
                         int32_t *source) {
  if (!variant) return fail(HML_STATUS_INVALID_VALUE, "variant is NULL");
  if (M <= 0 || N <= 0) return fail(HML_STATUS_INVALID_VALUE, "dimensions must be positive");
  GemvCanon c = canonicalize(order, trans, M, N, nullptr, 1, nullptr, nullptr);
  GemvKey key{c.family, c.outputs, c.reduced, beta == 0.0f ? 1 : 0};
  int src = 0;
  select_sgemv(key, variant, &src);
  if (source) *source = src;
  return HML_STATUS_SUCCESS;
}

void hmlSgemvBasic(float *y, const float *A, const float *x, const int M,
                   const int N, const float alpha, const float beta) {
  sgemv_compat(y, A, x, M, N, alpha, beta, true);
}

void hmlSgemv(float *y, const float *A, const float *x, const int M,
              const int N, const float alpha, const float beta) {
  sgemv_compat(y, A, x, M, N, alpha, beta, false);
}

}  // extern "C"
