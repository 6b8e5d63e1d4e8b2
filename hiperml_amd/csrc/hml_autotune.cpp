// hml_autotune.cpp -- exhaustive, validating autotuner and timing database.
//
// Counterpart of hmlSgemmKernelConfigOnline (src/hml_sgemm_kernel.cu:356-555)
// and hmlSgemvKernelConfigOnline (src/hml_sgemv_kernel.cu:391-632).  What
// stays the same: every candidate of the variant space is run `trials`
// times, the median decides, the winner per shape key becomes the config
// rule, a launch failure invalidates only that candidate.  What changes:
//  * every candidate is VALIDATED before it is timed (the reference times
//    only, which let its BasicN reduction win although it drops partial
//    sums for blockStops 6/7/9/10, include/hml_sgemv_kernel_template.h:67-100,
//    src/hml_sgemv_kernel.cu:491-533): the
//    shape's reference output is itself checked against the FP64 kernel on
//    three row windows, then each candidate's full output is compared with
//    it on the device;
//  * timing uses hipEvents on the stream around the kernels only, with an
//    optional 512 MiB cache flush before each trial so that the 256 MB MALL
//    does not flatter small shapes; inputs are uniform[-1,1] (the reference
//    times on memset byte patterns 11/22/33);
//  * every (shape, variant, median) is appended to a CSV timing database,
//    the training set of the learned selector; the database is also the
//    checkpoint (the reference's is the .cfg it writes at the end,
//    src/hml_sgemm_kernel.cu:557-602): a candidate already recorded for a
//    shape is not measured again, so an interrupted sweep resumes where it
//    stopped and a finished shape costs no GPU work;
//  * a fake timer can replace the GPU so the search/selection logic is
//    testable on a CPU.
#include "hml_autotune.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "hml_config.h"
#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_mlp.h"
#include "hml_select.h"

namespace hml {

namespace {

hmlAutotuneTimer g_fake_timer = nullptr;
void *g_fake_user = nullptr;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hmlStatus alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(HML_STATUS_ALLOC_FAILED, "autotune hipMalloc(%zu): %s", bytes,
                  hipGetErrorString(e));
    }
    return HML_STATUS_SUCCESS;
  }
  float *f() const { return (float *)p; }
};

struct Events {
  hipEvent_t a = nullptr, b = nullptr;
  ~Events() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  hmlStatus init() {
    hipError_t e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    return from_hip(e, "hipEventCreate");
  }
};

void dense_lds(const hmlGemmShape &s, int64_t *lda, int64_t *ldb, int64_t *ldc) {
  const bool rm = s.order == HML_ROW_MAJOR;
  *lda = std::max<int64_t>(1, rm ? (s.transA == HML_OP_N ? s.K : s.M) : (s.transA == HML_OP_N ? s.M : s.K));
  *ldb = std::max<int64_t>(1, rm ? (s.transB == HML_OP_N ? s.N : s.K) : (s.transB == HML_OP_N ? s.K : s.N));
  *ldc = std::max<int64_t>(1, rm ? s.N : s.M);
}

GemmKey shape_key(const hmlGemmShape &s) {
  int64_t lda, ldb, ldc;
  dense_lds(s, &lda, &ldb, &ldc);
  GemmCanon g = canonicalize((hmlOrder)s.order, (hmlOp)s.transA, (hmlOp)s.transB,
                             s.M, s.N, s.K, nullptr, lda, nullptr, ldb, nullptr, ldc);
  return make_key(g, s.beta0 ? 0.0f : 1.0f, 0, (hmlOrder)s.order,
                  (hmlOp)s.transA, (hmlOp)s.transB);
}

void push_unique(std::vector<hmlSgemmVariant> *out, const hmlSgemmVariant &v) {
  for (const hmlSgemmVariant &u : *out)
    if (u.family == v.family && u.rt == v.rt && u.ct == v.ct &&
        u.ksplit == v.ksplit && u.wpb == v.wpb)
      return;
  out->push_back(v);
}

void push_unique_v(std::vector<hmlSgemvVariant> *out, const hmlSgemvVariant &v) {
  for (const hmlSgemvVariant &u : *out)
    if (u.family == v.family && u.lpr == v.lpr && u.rb == v.rb &&
        u.split == v.split && u.wpb == v.wpb)
      return;
  out->push_back(v);
}

std::vector<int> split_ladder(int64_t heur, int64_t cap) {
  std::set<int> s;
  s.insert(1);
  for (int64_t f : {heur / 4, heur / 2, heur, heur * 2, heur * 4})
    if (f >= 1 && f <= cap) s.insert((int)f);
  return std::vector<int>(s.begin(), s.end());
}

}  // namespace

std::vector<hmlSgemvVariant> sgemv_candidates(const GemvKey &k) {
  std::vector<hmlSgemvVariant> out;
  hmlSgemvVariant h;
  heuristic_sgemv(k, &h);
  const int lens_lpr = h.lpr;
  for (int lpr = 4; lpr <= 64; lpr <<= 1) {
    if (lpr < lens_lpr / 4 || lpr > lens_lpr * 2) continue;
    const int64_t red = k.reduced;
    for (int rb : k.family == HML_SGEMV_FAMILY_DOT ? std::vector<int>{1, 2, 4, 8} : std::vector<int>{1}) {
      for (int wpb : {2, 4, 8}) {
        const int64_t cap = k.family == HML_SGEMV_FAMILY_DOT
                                ? std::max<int64_t>(1, red / (4 * lpr * 4))
                                : std::max<int64_t>(1, red / 64);
        for (int sp : split_ladder(h.split, std::min<int64_t>(cap, 4096))) {
          hmlSgemvVariant v{k.family, lpr, rb, sp, wpb};
          if (sgemv_variant_runnable(k, v)) push_unique_v(&out, v);
        }
      }
    }
  }
  if (sgemv_variant_runnable(k, h)) push_unique_v(&out, h);
  return out;
}

std::vector<hmlSgemmVariant> sgemm_candidates(const GemmKey &k) {
  std::vector<hmlSgemmVariant> out;
  hmlSgemmVariant h;
  heuristic_sgemm(k, &h);
  if (k.gemv_ok) {
    GemvKey gk{k.form == kFormDot ? 1 : 2, k.M, k.K, k.beta0};
    for (const hmlSgemvVariant &g : sgemv_candidates(gk))
      push_unique(&out, hmlSgemmVariant{HML_SGEMM_FAMILY_GEMV, g.lpr, g.rb, g.split, g.wpb});
  }
  // Every skinny / dense class of the learned selector that can run the
  // shape (row tiles 1/2/4, full or half column tiles, K split 1..64 up to
  // K/128, 2 or 4 waves per block), so the timing database labels exactly
  // the choices the selector can make.
  for (int c = 0; c < kSelClasses; ++c) {
    hmlSgemmVariant v;
    if (selector_variant(k, c, &v)) push_unique(&out, v);
  }
  // N = 2, 4, 8 on the multi-vector GEMV kernels: the DOT SGEMV parameter
  // sweep of the same rows (lanes per row 16..64, 1..4 rows per group).
  if (k.form == kFormDot && (k.N == 2 || k.N == 4 || k.N == 8)) {
    GemvKey gk{HML_SGEMV_FAMILY_DOT, k.M, k.K, k.beta0};
    for (const hmlSgemvVariant &g : sgemv_candidates(gk)) {
      hmlSgemmVariant v{HML_SGEMM_FAMILY_GEMVN, g.lpr, g.rb, g.split, g.wpb};
      if (sgemm_variant_runnable(k, v)) push_unique(&out, v);
    }
  }
  // Shared-B DOT tiles for wide column blocks (N > 32), K split 1..16.
  if (k.form == kFormDot && k.N > 32) {
    for (int rt : {1, 2})
      for (int ks : {1, 2, 4, 8, 16}) {
        hmlSgemmVariant v{HML_SGEMM_FAMILY_DOTB, rt, 4, ks, 4};
        if (ks <= std::max<int64_t>(1, k.K / 256) && sgemm_variant_runnable(k, v)) push_unique(&out, v);
      }
  }
  if (sgemm_variant_runnable(k, h)) push_unique(&out, h);
  return out;
}

namespace {

double median(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? -1.0 : v[v.size() / 2];
}

void db_write(FILE *db, int kind, const int64_t *dims, const int32_t *ctx,
              int form, const int32_t *var, double ms, int valid) {
  if (!db) return;
  fprintf(db, "%d,%lld,%lld,%lld,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%.6f,%d\n", kind,
          (long long)dims[0], (long long)dims[1], (long long)dims[2], ctx[0],
          ctx[1], ctx[2], ctx[3], form, var[0], var[1], var[2], var[3], var[4],
          ms, valid);
}

FILE *db_open(const char *path) {
  if (!path) return nullptr;
  FILE *f = fopen(path, "a+");
  if (!f) return nullptr;
  fseek(f, 0, SEEK_END);
  if (ftell(f) == 0)
    fprintf(f, "kind,M,N,K,order,transA,transB,beta0,form,family,p1,p2,p3,p4,median_ms,valid\n");
  return f;
}

// ---- timing database index (resume) ----------------------------------------
// Rows already in a database, by shape: the file is read incrementally (only
// what was appended since the last call), so a sweep over thousands of
// shapes does not re-read it per shape.
struct DbRow {
  int32_t var[5];
  double ms;
  int valid;
};
struct DbIndex {
  std::string path;
  long offset = 0;
  std::map<std::string, std::vector<DbRow>> rows;
};
std::mutex g_db_mutex;
DbIndex g_db;

std::string db_key(int kind, const int64_t *dims, const int32_t *ctx) {
  char b[160];
  snprintf(b, sizeof b, "%d,%lld,%lld,%lld,%d,%d,%d,%d", kind, (long long)dims[0],
           (long long)dims[1], (long long)dims[2], ctx[0], ctx[1], ctx[2], ctx[3]);
  return b;
}

// Rows recorded for one shape (empty without a database).
std::vector<DbRow> db_prior(const char *path, int kind, const int64_t *dims,
                            const int32_t *ctx) {
  std::vector<DbRow> out;
  if (!path) return out;
  std::lock_guard<std::mutex> lk(g_db_mutex);
  if (g_db.path != path) {
    g_db = DbIndex();
    g_db.path = path;
  }
  FILE *f = fopen(path, "r");
  if (f) {
    fseek(f, 0, SEEK_END);
    const long end = ftell(f);
    if (end < g_db.offset) {  // the file was replaced: start over
      g_db.rows.clear();
      g_db.offset = 0;
    }
    fseek(f, g_db.offset, SEEK_SET);
    char line[512];
    while (fgets(line, sizeof line, f)) {
      const size_t len = strlen(line);
      if (len == 0 || line[len - 1] != '\n') break;  // partial last line
      g_db.offset += (long)len;
      int k, c[4], form, v[5], valid;
      long long d[3];
      double ms;
      if (sscanf(line, "%d,%lld,%lld,%lld,%d,%d,%d,%d,%d,%d,%d,%d,%d,%d,%lf,%d", &k, &d[0],
                 &d[1], &d[2], &c[0], &c[1], &c[2], &c[3], &form, &v[0], &v[1], &v[2],
                 &v[3], &v[4], &ms, &valid) != 16)
        continue;  // header or foreign line
      const int64_t dd[3] = {d[0], d[1], d[2]};
      const int32_t cc[4] = {c[0], c[1], c[2], c[3]};
      g_db.rows[db_key(k, dd, cc)].push_back(DbRow{{v[0], v[1], v[2], v[3], v[4]}, ms, valid});
    }
    fclose(f);
  }
  auto it = g_db.rows.find(db_key(kind, dims, ctx));
  if (it != g_db.rows.end()) out = it->second;
  return out;
}

// The recorded row of one variant (the latest, if recorded twice).
const DbRow *db_find(const std::vector<DbRow> &rows, const int32_t *var) {
  const DbRow *hit = nullptr;
  for (const DbRow &r : rows)
    if (!memcmp(r.var, var, sizeof r.var)) hit = &r;
  return hit;
}

// FP64 check of rows [m0, m0+mw) of an SGEMM output.
hmlStatus check_window(const hmlGemmShape &s, int64_t lda, int64_t ldb,
                       int64_t ldc, const float *A, const float *B,
                       const float *C0, const float *C, int64_t m0, int64_t mw,
                       double *st6) {
  const bool rm = s.order == HML_ROW_MAJOR;
  const int64_t sAm = (rm == (s.transA == HML_OP_N)) ? lda : 1;
  const int64_t sCm = rm ? ldc : 1;
  return hmlSgemmCheckF64((hmlOrder)s.order, (hmlOp)s.transA, (hmlOp)s.transB,
                          mw, s.N, s.K, 1.0f, A + m0 * sAm, lda, B, ldb,
                          s.beta0 ? 0.0f : 0.5f, s.beta0 ? nullptr : C0 + m0 * sCm,
                          C + m0 * sCm, ldc, st6);
}

hmlStatus compare_dev(const float *C, const float *R, int64_t M, int64_t N,
                      int64_t sCm, int64_t sCn, double *dstats, double *out5) {
  hipError_t e = launch_compare(C, R, M, N, sCm, sCn, dstats, kCompareBlocks, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return from_hip(e, "compare");
  std::vector<double> h((size_t)kCompareBlocks * 5);
  e = hipMemcpy(h.data(), dstats, h.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return from_hip(e, "compare copy");
  for (int i = 0; i < 5; ++i) out5[i] = 0.0;
  for (int b = 0; b < kCompareBlocks; ++b) {
    out5[0] += h[b * 5 + 0];
    out5[1] += h[b * 5 + 1];
    out5[2] = std::max(out5[2], h[b * 5 + 2]);
    out5[3] = std::max(out5[3], h[b * 5 + 3]);
    out5[4] += h[b * 5 + 4];
  }
  return HML_STATUS_SUCCESS;
}

// Relative-Frobenius acceptance threshold between two FP32 results of a
// K-term reduction on uniform[-1,1] data: genuine rounding differences
// are O(sqrt(K) u); an indexing or tiling error is O(1).
double accept_threshold(int64_t K) {
  return 1e-4 + 8.0 * sqrt((double)std::max<int64_t>(K, 1)) * ldexp(1.0, -24);
}

}  // namespace

// ---- summary tables ------------------------------------------------------
// The reference prints, after its online sweep, per kernel type a table of
// the best GFLOPS over row tiles for every (K, N) and a table of max / min
// GFLOPS over them (hmlSgemmKernelConfigPrintStats,
// src/hml_sgemm_kernel.cu:318-354; the SGEMV sweep prints the same per N,
// src/hml_sgemv_kernel.cu:391-632).  Here the tables come from a timing
// database: per layout group (SGEMM: M, order, transA, transB, beta == 0;
// SGEMV: order, trans) and per family, rows K (SGEMV: M) and columns N, the
// best GFLOP/s over the family's variants and the spread max / min of its
// valid variants in percent; a last table gives the overall winner's
// family and GFLOP/s.
const char *family_name(int kind, int f) {
  static const char *gemm[] = {"basic", "dot", "axpy", "gemv", "bf16x3",
                               "dense", "bf16x6", "dotb", "gemvn"};
  static const char *gemv[] = {"basic", "dot", "axpy"};
  if (kind == 0) return f >= 0 && f < 9 ? gemm[f] : "?";
  return f >= 0 && f < 3 ? gemv[f] : "?";
}

void print_stats(const char *db_path, int kind, FILE *out) {
  const int64_t none[3] = {0, 0, 0};
  const int32_t nctx[4] = {0, 0, 0, 0};
  (void)db_prior(db_path, kind, none, nctx);  // refresh the index
  struct Cell {
    double best = 0.0, lo = 0.0, hi = 0.0;
  };
  // group -> family -> (row, col) -> cell; group -> (row, col) -> (best, family)
  std::map<std::string, std::map<int, std::map<std::pair<int64_t, int64_t>, Cell>>> tab;
  std::map<std::string, std::map<std::pair<int64_t, int64_t>, std::pair<double, int>>> win;
  std::lock_guard<std::mutex> lk(g_db_mutex);
  for (const auto &kv : g_db.rows) {
    int k, c[4];
    long long d[3];
    if (sscanf(kv.first.c_str(), "%d,%lld,%lld,%lld,%d,%d,%d,%d", &k, &d[0], &d[1], &d[2],
               &c[0], &c[1], &c[2], &c[3]) != 8 || k != kind)
      continue;
    char g[96];
    int64_t row, col;
    double flops;
    if (kind == 0) {
      snprintf(g, sizeof g, "M=%lld order=%d transA=%d transB=%d beta0=%d", d[0], c[0], c[1],
               c[2], c[3]);
      row = d[2]; col = d[1];
      flops = 2.0 * d[0] * d[1] * d[2];
    } else {
      snprintf(g, sizeof g, "order=%d trans=%d", c[0], c[1]);
      row = d[0]; col = d[1];
      flops = 2.0 * d[0] * d[1] + 2.0 * d[0];
    }
    for (const DbRow &r : kv.second) {
      if (!r.valid || r.ms <= 0) continue;
      const double gf = flops / (r.ms * 1e6);
      Cell &cell = tab[g][r.var[0]][{row, col}];
      if (cell.best == 0.0) cell.lo = gf;
      cell.best = std::max(cell.best, gf);
      cell.hi = std::max(cell.hi, gf);
      cell.lo = std::min(cell.lo, gf);
      auto &w = win[g][{row, col}];
      if (gf > w.first) w = {gf, r.var[0]};
    }
  }
  const char *rowname = kind == 0 ? "K" : "M";
  for (const auto &gt : tab) {
    std::set<int64_t> rows, cols;
    for (const auto &ft : gt.second)
      for (const auto &c : ft.second) { rows.insert(c.first.first); cols.insert(c.first.second); }
    auto header = [&](const char *title) {
      fprintf(out, "%s [%s]:\n%8s", title, gt.first.c_str(), rowname);
      for (int64_t n : cols) fprintf(out, " %9s", ("N=" + std::to_string(n)).c_str());
      fprintf(out, "\n");
      for (size_t i = 0; i < 8 + 10 * cols.size(); ++i) fputc('=', out);
      fprintf(out, "\n");
    };
    for (const auto &ft : gt.second) {
      char title[64];
      snprintf(title, sizeof title, "%s GFLOPS", family_name(kind, ft.first));
      header(title);
      for (int64_t r : rows) {
        fprintf(out, "%8lld", (long long)r);
        for (int64_t n : cols) {
          auto it = ft.second.find({r, n});
          if (it == ft.second.end()) fprintf(out, " %9s", "-");
          else fprintf(out, " %9.2f", it->second.best);
        }
        fprintf(out, "\n");
      }
      snprintf(title, sizeof title, "%s max / min", family_name(kind, ft.first));
      header(title);
      for (int64_t r : rows) {
        fprintf(out, "%8lld", (long long)r);
        for (int64_t n : cols) {
          auto it = ft.second.find({r, n});
          if (it == ft.second.end() || it->second.lo <= 0) fprintf(out, " %9s", "-");
          else fprintf(out, " %8.0f%%", it->second.hi * 100.0 / it->second.lo);
        }
        fprintf(out, "\n");
      }
    }
    header("best GFLOPS (family)");
    for (int64_t r : rows) {
      fprintf(out, "%8lld", (long long)r);
      for (int64_t n : cols) {
        auto it = win[gt.first].find({r, n});
        if (it == win[gt.first].end()) {
          fprintf(out, " %9s", "-");
        } else {
          char cell[32];
          snprintf(cell, sizeof cell, "%.1f(%d)", it->second.first, it->second.second);
          fprintf(out, " %9s", cell);
        }
      }
      fprintf(out, "\n");
    }
  }
}

hmlStatus autotune_sgemm_shape(const hmlGemmShape &s, int trials, int flush,
                               uint64_t seed, hmlSgemmVariant *best,
                               double *best_ms, double *heur_ms, FILE *db,
                               const char *db_path, int verbose) {
  if (s.M <= 0 || s.N <= 0 || s.K <= 0)
    return fail(HML_STATUS_INVALID_VALUE, "autotune shape must be positive");
  if (trials < 1) trials = 1;
  const GemmKey key = shape_key(s);
  const std::vector<hmlSgemmVariant> cands = sgemm_candidates(key);
  hmlSgemmVariant h;
  heuristic_sgemm(key, &h);
  const int64_t dims[3] = {s.M, s.N, s.K};
  const int32_t ctx[4] = {s.order, s.transA, s.transB, s.beta0};
  double bestt = -1.0, heurt = -1.0;
  hmlSgemmVariant bestv = h;
  auto same = [](const hmlSgemmVariant &a, const hmlSgemmVariant &b) {
    return a.family == b.family && a.rt == b.rt && a.ct == b.ct && a.ksplit == b.ksplit &&
           a.wpb == b.wpb;
  };
  // Resume: candidates recorded by an earlier run count with their recorded
  // median (or failure) and are not measured again.
  const std::vector<DbRow> prior = db_prior(db_path, 0, dims, ctx);
  std::vector<hmlSgemmVariant> todo;
  for (const hmlSgemmVariant &v : cands) {
    const int32_t var[5] = {v.family, v.rt, v.ct, v.ksplit, v.wpb};
    const DbRow *r = db_find(prior, var);
    if (!r) {
      todo.push_back(v);
      continue;
    }
    const double t = r->valid ? r->ms : -1.0;
    if (t < 0) continue;
    if (same(v, h)) heurt = t;
    if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
  }
  auto finish = [&]() -> hmlStatus {
    if (bestt < 0) return fail(HML_STATUS_VALIDATION_FAILED, "no candidate passed validation");
    *best = bestv;
    if (best_ms) *best_ms = bestt;
    if (heur_ms) *heur_ms = heurt;
    return HML_STATUS_SUCCESS;
  };
  if (todo.empty()) return finish();

  if (g_fake_timer) {
    for (const hmlSgemmVariant &v : todo) {
      const int32_t var[5] = {v.family, v.rt, v.ct, v.ksplit, v.wpb};
      const double t = g_fake_timer(0, dims, var, g_fake_user);
      db_write(db, 0, dims, ctx, key.form, var, t, t >= 0 ? 1 : 0);
      if (t < 0) continue;
      if (same(v, h)) heurt = t;
      if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
    }
    return finish();
  }

  int64_t lda, ldb, ldc;
  dense_lds(s, &lda, &ldb, &ldc);
  const bool rm = s.order == HML_ROW_MAJOR;
  const size_t nA = (size_t)lda * (rm == (s.transA == HML_OP_N) ? s.M : s.K);
  const size_t nB = (size_t)ldb * (rm == (s.transB == HML_OP_N) ? s.K : s.N);
  const size_t nC = (size_t)ldc * (rm ? s.M : s.N);
  size_t ws = 0;
  for (const hmlSgemmVariant &v : cands)
    ws = std::max(ws, hmlSgemmWorkspaceSize((hmlOrder)s.order, (hmlOp)s.transA,
                                            (hmlOp)s.transB, s.M, s.N, s.K, &v));
  DevBuf A, B, C0, C, R, W, S;
  hmlStatus st;
  if ((st = A.alloc(nA * 4)) || (st = B.alloc(nB * 4)) || (st = C0.alloc(nC * 4)) ||
      (st = C.alloc(nC * 4)) || (st = R.alloc(nC * 4)) || (st = W.alloc(ws)) ||
      (st = S.alloc((size_t)kCompareBlocks * 5 * sizeof(double))))
    return st;
  hipError_t e = launch_fill_uniform(A.f(), nA, seed * 3 + 1, -1.f, 1.f, nullptr);
  if (e == hipSuccess) e = launch_fill_uniform(B.f(), nB, seed * 3 + 2, -1.f, 1.f, nullptr);
  if (e == hipSuccess) e = launch_fill_uniform(C0.f(), nC, seed * 3 + 3, -1.f, 1.f, nullptr);
  if (e != hipSuccess) return from_hip(e, "autotune fill");
  const float alpha = 1.0f, beta = s.beta0 ? 0.0f : 0.5f;
  const int64_t sCm = rm ? ldc : 1, sCn = rm ? 1 : ldc;
  Events ev;
  if ((st = ev.init())) return st;

  auto run = [&](const hmlSgemmVariant &v, float *Cout) -> hmlStatus {
    return hmlSgemmExVariant((hmlOrder)s.order, (hmlOp)s.transA, (hmlOp)s.transB,
                             s.M, s.N, s.K, alpha, A.f(), lda, B.f(), ldb, beta,
                             Cout, ldc, 0, &v, W.p, ws, nullptr);
  };

  // Reference output: the heuristic variant, FP64-checked on 3 windows.
  e = hipMemcpy(R.p, C0.p, nC * 4, hipMemcpyDeviceToDevice);
  if (e != hipSuccess) return from_hip(e, "autotune copy");
  if ((st = run(h, R.f()))) return st;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return from_hip(e, "reference run");
  {
    const int64_t mw = std::min<int64_t>(s.M, 64);
    const int64_t starts[3] = {0, (s.M - mw) / 2, s.M - mw};
    for (int w = 0; w < 3; ++w) {
      double st6[6];
      if ((st = check_window(s, lda, ldb, ldc, A.f(), B.f(), C0.f(), R.f(), starts[w], mw, st6)))
        return st;
      const double u = ldexp(1.0, -24);
      const double gam = (double)(s.K + 2) * u / (1.0 - (double)(s.K + 2) * u);
      if (st6[5] > 0 || sqrt(st6[0]) > gam * sqrt(st6[2]) + 1e-30)
        return fail(HML_STATUS_VALIDATION_FAILED,
                    "reference variant failed the FP64 check (err %.3g, bound %.3g)",
                    sqrt(st6[0]), gam * sqrt(st6[2]));
    }
  }
  const double thr = accept_threshold(s.K);
  for (const hmlSgemmVariant &v : todo) {
    const int32_t var[5] = {v.family, v.rt, v.ct, v.ksplit, v.wpb};
    e = hipMemcpy(C.p, C0.p, nC * 4, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) return from_hip(e, "autotune copy");
    st = run(v, C.f());
    if (st == HML_STATUS_SUCCESS) {
      e = hipDeviceSynchronize();
      if (e != hipSuccess) return from_hip(e, "candidate run");  // device error is fatal
    }
    int valid = 0;
    if (st == HML_STATUS_SUCCESS) {
      double c5[5];
      if ((st = compare_dev(C.f(), R.f(), s.M, s.N, sCm, sCn, S.p ? (double *)S.p : nullptr, c5)))
        return st;
      valid = c5[4] == 0 && sqrt(c5[0]) <= thr * sqrt(c5[1]) + 1e-30;
    }
    double t = -1.0;
    if (valid) {
      std::vector<double> ts;
      for (int i = 0; i < trials; ++i) {
        if (flush && hmlFlushCaches(nullptr) != HML_STATUS_SUCCESS) return HML_STATUS_HIP_ERROR;
        if ((e = hipEventRecord(ev.a, nullptr)) != hipSuccess) return from_hip(e, "timing");
        if ((st = run(v, C.f()))) break;
        if ((e = hipEventRecord(ev.b, nullptr)) != hipSuccess) return from_hip(e, "timing");
        if ((e = hipEventSynchronize(ev.b)) != hipSuccess) return from_hip(e, "timing");
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, ev.a, ev.b)) != hipSuccess) return from_hip(e, "timing");
        ts.push_back(ms);
      }
      if (st == HML_STATUS_SUCCESS) t = median(ts);
    } else if (verbose) {
      fprintf(stderr, "Warning: variant [%d %d %d %d %d] failed validation for M=%lld N=%lld K=%lld\n",
              v.family, v.rt, v.ct, v.ksplit, v.wpb, (long long)s.M, (long long)s.N, (long long)s.K);
    }
    db_write(db, 0, dims, ctx, key.form, var, t, t >= 0 ? 1 : 0);
    if (db) fflush(db);
    if (t < 0) continue;
    if (same(v, h)) heurt = t;
    if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
    if (verbose >= 2)
      fprintf(stderr, "Info: kernel[%lld][%lld][%lld][%d %d %d %d %d] = %.4f ms\n",
              (long long)s.M, (long long)s.N, (long long)s.K, v.family, v.rt,
              v.ct, v.ksplit, v.wpb, t);
  }
  return finish();
}

hmlStatus autotune_sgemv_shape(hmlOrder order, hmlOp trans, int64_t M,
                               int64_t N, int trials, int flush, uint64_t seed,
                               hmlSgemvVariant *best, double *best_ms,
                               double *heur_ms, FILE *db, const char *db_path,
                               int verbose) {
  if (M <= 0 || N <= 0) return fail(HML_STATUS_INVALID_VALUE, "autotune shape must be positive");
  if (trials < 1) trials = 1;
  const GemvKey key = gemv_key(order, trans, M, N, 1);
  const std::vector<hmlSgemvVariant> cands = sgemv_candidates(key);
  hmlSgemvVariant h;
  heuristic_sgemv(key, &h);
  const int64_t dims[3] = {M, N, 0};
  const int32_t ctx[4] = {order, trans, 0, 1};
  double bestt = -1.0, heurt = -1.0;
  hmlSgemvVariant bestv = h;
  auto same = [](const hmlSgemvVariant &a, const hmlSgemvVariant &b) {
    return a.family == b.family && a.lpr == b.lpr && a.rb == b.rb && a.split == b.split && a.wpb == b.wpb;
  };
  const std::vector<DbRow> prior = db_prior(db_path, 1, dims, ctx);
  std::vector<hmlSgemvVariant> todo;
  for (const hmlSgemvVariant &v : cands) {
    const int32_t var[5] = {v.family, v.lpr, v.rb, v.split, v.wpb};
    const DbRow *r = db_find(prior, var);
    if (!r) {
      todo.push_back(v);
      continue;
    }
    const double t = r->valid ? r->ms : -1.0;
    if (t < 0) continue;
    if (same(v, h)) heurt = t;
    if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
  }
  auto finish = [&]() -> hmlStatus {
    if (bestt < 0) return fail(HML_STATUS_VALIDATION_FAILED, "no candidate passed validation");
    *best = bestv;
    if (best_ms) *best_ms = bestt;
    if (heur_ms) *heur_ms = heurt;
    return HML_STATUS_SUCCESS;
  };
  if (todo.empty()) return finish();
  if (g_fake_timer) {
    for (const hmlSgemvVariant &v : todo) {
      const int32_t var[5] = {v.family, v.lpr, v.rb, v.split, v.wpb};
      const double t = g_fake_timer(1, dims, var, g_fake_user);
      db_write(db, 1, dims, ctx, key.family, var, t, t >= 0 ? 1 : 0);
      if (t < 0) continue;
      if (same(v, h)) heurt = t;
      if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
    }
    return finish();
  }
  const int64_t lda = order == HML_ROW_MAJOR ? N : M;
  const size_t nA = (size_t)M * N;
  const int64_t nx = trans == HML_OP_N ? N : M, ny = trans == HML_OP_N ? M : N;
  size_t ws = 0;
  for (const hmlSgemvVariant &v : cands)
    ws = std::max(ws, hmlSgemvWorkspaceSize(order, trans, M, N, &v));
  DevBuf A, x, y, R, W, S, y64;
  hmlStatus st;
  if ((st = A.alloc(nA * 4)) || (st = x.alloc(nx * 4)) || (st = y.alloc(ny * 4)) ||
      (st = R.alloc(ny * 4)) || (st = W.alloc(ws)) ||
      (st = S.alloc((size_t)kCompareBlocks * 5 * sizeof(double))))
    return st;
  hipError_t e = launch_fill_uniform(A.f(), nA, seed * 3 + 1, -1.f, 1.f, nullptr);
  if (e == hipSuccess) e = launch_fill_uniform(x.f(), nx, seed * 3 + 2, -1.f, 1.f, nullptr);
  if (e != hipSuccess) return from_hip(e, "autotune fill");
  Events ev;
  if ((st = ev.init())) return st;
  auto run = [&](const hmlSgemvVariant &v, float *yout) -> hmlStatus {
    return hmlSgemvExVariant(order, trans, M, N, 1.0f, A.f(), lda, x.f(), 1, 0.0f,
                             yout, 1, 0, &v, W.p, ws, nullptr);
  };
  if ((st = run(h, R.f()))) return st;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return from_hip(e, "reference run");
  {
    // FP64 check of the reference output: y = op(A) x as a GEMM with N = 1.
    double st6[6];
    const hmlOp tb = HML_OP_N;
    const int64_t ldx = order == HML_ROW_MAJOR ? 1 : nx;
    const int64_t ldy = order == HML_ROW_MAJOR ? 1 : ny;
    if ((st = hmlSgemmCheckF64(order, trans, tb, ny, 1, nx, 1.0f, A.f(), lda,
                               x.f(), ldx, 0.0f, nullptr, R.f(), ldy, st6)))
      return st;
    const double u = ldexp(1.0, -24);
    const double gam = (double)(nx + 2) * u / (1.0 - (double)(nx + 2) * u);
    if (st6[5] > 0 || sqrt(st6[0]) > gam * sqrt(st6[2]) + 1e-30)
      return fail(HML_STATUS_VALIDATION_FAILED, "reference GEMV failed the FP64 check");
  }
  const double thr = accept_threshold(nx);
  for (const hmlSgemvVariant &v : todo) {
    const int32_t var[5] = {v.family, v.lpr, v.rb, v.split, v.wpb};
    st = run(v, y.f());
    if (st == HML_STATUS_SUCCESS && (e = hipDeviceSynchronize()) != hipSuccess)
      return from_hip(e, "candidate run");
    int valid = 0;
    if (st == HML_STATUS_SUCCESS) {
      double c5[5];
      if ((st = compare_dev(y.f(), R.f(), ny, 1, 1, 0, (double *)S.p, c5))) return st;
      valid = c5[4] == 0 && sqrt(c5[0]) <= thr * sqrt(c5[1]) + 1e-30;
    }
    double t = -1.0;
    if (valid) {
      std::vector<double> ts;
      for (int i = 0; i < trials; ++i) {
        if (flush && hmlFlushCaches(nullptr) != HML_STATUS_SUCCESS) return HML_STATUS_HIP_ERROR;
        if ((e = hipEventRecord(ev.a, nullptr)) != hipSuccess) return from_hip(e, "timing");
        if ((st = run(v, y.f()))) break;
        if ((e = hipEventRecord(ev.b, nullptr)) != hipSuccess) return from_hip(e, "timing");
        if ((e = hipEventSynchronize(ev.b)) != hipSuccess) return from_hip(e, "timing");
        float ms = 0.f;
        if ((e = hipEventElapsedTime(&ms, ev.a, ev.b)) != hipSuccess) return from_hip(e, "timing");
        ts.push_back(ms);
      }
      if (st == HML_STATUS_SUCCESS) t = median(ts);
    } else if (verbose) {
      fprintf(stderr, "Warning: gemv variant [%d %d %d %d %d] failed validation\n",
              v.family, v.lpr, v.rb, v.split, v.wpb);
    }
    db_write(db, 1, dims, ctx, key.family, var, t, t >= 0 ? 1 : 0);
    if (db) fflush(db);
    if (t < 0) continue;
    if (same(v, h)) heurt = t;
    if (bestt < 0 || t < bestt) { bestt = t; bestv = v; }
  }
  (void)y64;
  return finish();
}

// Database of the online sweeps: $HML_AUTOTUNE_DB when set (the sweep then
// resumes from it and appends to it), else none.
const char *online_db_path() {
  const char *p = getenv("HML_AUTOTUNE_DB");
  return p && *p ? p : nullptr;
}

hmlStatus autotune_sgemm_online(int verbose) {
  // Row-major NN grid (the layout of the reference API): M 2^12..2^20,
  // N 1..128, K 2^5..2^13, at most 1 GiB per operand.
  const char *db_path = online_db_path();
  FILE *db = db_open(db_path);
  std::vector<SgemmRuleV2> rules;
  for (int ml = 12; ml <= 20; ml += 2)
    for (int nl = 0; nl <= 7; ++nl)
      for (int kl = 5; kl <= 13; kl += 2) {
        hmlGemmShape s{(int64_t)1 << ml, (int64_t)1 << nl, (int64_t)1 << kl,
                       HML_ROW_MAJOR, HML_OP_N, HML_OP_N, 1};
        if ((double)s.M * s.K * 4 > (double)(1 << 30)) continue;
        hmlSgemmVariant best;
        double bt = 0, ht = 0;
        hmlStatus st = autotune_sgemm_shape(s, 5, 1, 1, &best, &bt, &ht, db, db_path, verbose);
        if (st != HML_STATUS_SUCCESS) {
          if (db) fclose(db);
          return st;
        }
        const GemmKey k = shape_key(s);
        rules.push_back(SgemmRuleV2{k.form, ilog2_floor(k.M), ilog2_floor(k.N), ilog2_floor(k.K), best});
        if (verbose)
          fprintf(stderr, "Info: kernel*[M=%lld][N=%lld][K=%lld] = [%d %d %d %d %d] %.4f ms\n",
                  (long long)s.M, (long long)s.N, (long long)s.K, best.family,
                  best.rt, best.ct, best.ksplit, best.wpb, bt);
      }
  if (db) fclose(db);
  config_set_sgemm_v2(rules);
  if (verbose && db_path) print_stats(db_path, 0, stderr);
  return HML_STATUS_SUCCESS;
}

hmlStatus autotune_sgemv_online(int verbose) {
  const char *db_path = online_db_path();
  FILE *db = db_open(db_path);
  std::vector<SgemvRuleV2> rules;
  for (int trans = 0; trans <= 1; ++trans)
    for (int ml = 10; ml <= 20; ml += 2)
      for (int nl = 2; nl <= 14; nl += 2) {
        if (ml + nl > 28) continue;
        const int64_t M = (int64_t)1 << ml, N = (int64_t)1 << nl;
        hmlSgemvVariant best;
        double bt = 0, ht = 0;
        hmlStatus st = autotune_sgemv_shape(HML_ROW_MAJOR, (hmlOp)trans, M, N, 5, 1, 1,
                                            &best, &bt, &ht, db, db_path, verbose);
        if (st != HML_STATUS_SUCCESS) {
          if (db) fclose(db);
          return st;
        }
        const GemvKey k = gemv_key(HML_ROW_MAJOR, (hmlOp)trans, M, N, 1);
        rules.push_back(SgemvRuleV2{k.family, ilog2_floor(k.outputs), ilog2_floor(k.reduced), best});
        if (verbose)
          fprintf(stderr, "Info: kernel*[M=%lld][N=%lld][T=%d] = [%d %d %d %d %d] %.4f ms\n",
                  (long long)M, (long long)N, trans, best.family, best.lpr,
                  best.rb, best.split, best.wpb, bt);
      }
  if (db) fclose(db);
  config_set_sgemv_v2(rules);
  if (verbose && db_path) print_stats(db_path, 1, stderr);
  return HML_STATUS_SUCCESS;
}

}  // namespace hml

using namespace hml;

extern "C" {

void hmlAutotuneSetTimer(hmlAutotuneTimer fn, void *user) {
  g_fake_timer = fn;
  g_fake_user = user;
}

int32_t hmlSgemmCandidates(const hmlGemmShape *shape, hmlSgemmVariant *out,
                           int32_t cap) {
  if (!shape || shape->M <= 0 || shape->N <= 0 || shape->K <= 0) return -1;
  const std::vector<hmlSgemmVariant> c = sgemm_candidates(shape_key(*shape));
  for (int32_t i = 0; i < cap && i < (int32_t)c.size(); ++i) out[i] = c[i];
  return (int32_t)c.size();
}

int32_t hmlSgemvCandidates(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                           hmlSgemvVariant *out, int32_t cap) {
  if (M <= 0 || N <= 0) return -1;
  const std::vector<hmlSgemvVariant> c = sgemv_candidates(gemv_key(order, trans, M, N, 1));
  for (int32_t i = 0; i < cap && i < (int32_t)c.size(); ++i) out[i] = c[i];
  return (int32_t)c.size();
}

hmlStatus hmlSgemmAutotune(const hmlGemmShape *shape, int32_t trials,
                           int32_t flush, uint64_t seed,
                           hmlSgemmVariant *best, double *best_ms,
                           double *heuristic_ms, const char *db_path) {
  clear_error();
  if (!shape || !best) return fail(HML_STATUS_INVALID_VALUE, "shape/best is NULL");
  FILE *db = db_open(db_path);
  if (db_path && !db) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", db_path);
  hmlStatus s = autotune_sgemm_shape(*shape, trials, flush, seed, best, best_ms,
                                     heuristic_ms, db, db_path, verbosity() >= 2 ? 2 : 0);
  if (db) fclose(db);
  return s;
}

hmlStatus hmlSgemvAutotune(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                           int32_t trials, int32_t flush, uint64_t seed,
                           hmlSgemvVariant *best, double *best_ms,
                           double *heuristic_ms, const char *db_path) {
  clear_error();
  if (!best) return fail(HML_STATUS_INVALID_VALUE, "best is NULL");
  FILE *db = db_open(db_path);
  if (db_path && !db) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", db_path);
  hmlStatus s = autotune_sgemv_shape(order, trans, M, N, trials, flush, seed, best,
                                     best_ms, heuristic_ms, db, db_path,
                                     verbosity() >= 2 ? 2 : 0);
  if (db) fclose(db);
  return s;
}

hmlStatus hmlAutotuneStats(const char *db_path, int32_t kind, const char *out_path) {
  clear_error();
  if (!db_path || (kind != 0 && kind != 1))
    return fail(HML_STATUS_INVALID_VALUE, "autotune stats: db_path and kind 0 or 1 required");
  FILE *probe = fopen(db_path, "r");
  if (!probe) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", db_path);
  fclose(probe);
  FILE *out = out_path ? fopen(out_path, "w") : stderr;
  if (!out) return fail(HML_STATUS_IO_ERROR, "cannot create '%s'", out_path);
  print_stats(db_path, kind, out);
  if (out != stderr) fclose(out);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemmConfigAddRule(const hmlGemmShape *shape,
                                const hmlSgemmVariant *v) {
  if (!shape || !v || shape->M <= 0 || shape->N <= 0 || shape->K <= 0)
    return fail(HML_STATUS_INVALID_VALUE, "bad rule");
  const GemmKey k = shape_key(*shape);
  if (!sgemm_variant_runnable(k, *v))
    return fail(HML_STATUS_NO_KERNEL, "rule variant cannot run its shape");
  SgemmRuleV2 r{k.form, ilog2_floor(k.M), ilog2_floor(k.N), ilog2_floor(k.K), *v};
  config_add_sgemm_v2(r);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemvConfigAddRule(hmlOrder order, hmlOp trans, int64_t M,
                                int64_t N, const hmlSgemvVariant *v) {
  if (!v || M <= 0 || N <= 0) return fail(HML_STATUS_INVALID_VALUE, "bad rule");
  const GemvKey k = gemv_key(order, trans, M, N, 1);
  if (!sgemv_variant_runnable(k, *v))
    return fail(HML_STATUS_NO_KERNEL, "rule variant cannot run its shape");
  SgemvRuleV2 r{k.family, ilog2_floor(k.outputs), ilog2_floor(k.reduced), *v};
  config_add_sgemv_v2(r);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlFillUniform(float *dst, size_t n, uint64_t seed, float lo,
                         float hi, void *stream) {
  return from_hip(launch_fill_uniform(dst, n, seed, lo, hi, (hipStream_t)stream),
                  "fill uniform");
}

}  // extern "C"
