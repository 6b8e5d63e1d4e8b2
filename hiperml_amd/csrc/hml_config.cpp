// hml_config.cpp -- kernel-config store and table selection.
//
// v1.0 files are the reference's format and are read and written with the
// reference's grammar (an interface that must stay alike so the shipped
// per-GPU tables load unchanged):
//   SGEMM  "K colStops : kernelType rowStops"          (src/hml_sgemm_kernel.cu:18-36,557-602)
//   SGEMV  "N : kernelType blockStops rowStops"         (src/hml_sgemv_kernel.cu:19-37,634-679)
// Lines starting with '#' are comments; a rule with rowStops <= 0 (or
// blockStops <= 0) is invalid and selects the basic kernel
// (src/hml_sgemm_kernel.cu:216-220, src/hml_sgemv_kernel.cu:331-336).
// Unlike the reference reader, keys outside the table are skipped instead of
// written out of bounds, and lines with fewer than four fields are ignored.
//
// v2.0 (MI355X) files hold native variants keyed by power-of-two shape
// buckets, produced by the exhaustive autotuner:
//   SGEMM  "form log2M log2N log2K : family rt ct ksplit wpb"
//   SGEMV  "family log2outputs log2reduced : lpr rb split wpb"
// Lookup takes the exact bucket, else the nearest bucket of the same form
// (L1 distance in log2 space).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "hml_config.h"
#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_select.h"

namespace hml {

namespace {

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

std::mutex g_cfg_mutex;

// SGEMM v1 table, keys K 0..96 (cHmlMaxSkinnyK) x colStops 0..16
// (cHmlMaxStops), include/hml_sgemm_kernel.h:28-30.
SgemmRuleV1 g_sgemm_v1[kV1MaxSkinnyK + 1][kV1MaxStops + 1];
bool g_sgemm_v1_loaded = false;
// SGEMV v1 table, keys N 0..1024 (cHmlMaxSkinnyN), include/hml_sgemv_kernel.h:28.
SgemvRuleV1 g_sgemv_v1[kV1MaxSkinnyN + 1];
bool g_sgemv_v1_loaded = false;

std::vector<SgemmRuleV2> g_sgemm_v2;
std::vector<SgemvRuleV2> g_sgemv_v2;

bool g_sgemm_config_done = false;
bool g_sgemv_config_done = false;

const char kSgemmHeaderV1[] =
    "#SGEMM smart kernel config file v1.0\n"
    "#SGEMM: Single-precision general matrix-matrix multiplication\n"
    "#formula: C = alpha * A * B + beta * C\n"
    "#where:   A is M x K\n"
    "#         B is K x N\n"
    "#         C is M x N\n"
    "#         alpha and beta are scalars\n"
    "#format: K colStops : kernelType rowStops\n"
    "#where:\n"
    "#  K          : number of columns of A\n"
    "#  colStops   : (N + 15) / 16, number of column stops\n"
    "#  kernelType : {basic = 0, var-K = 1, const-K = 2}\n"
    "#  rowStops   : (#rows of A assigned to a 16x16 thread block) / 16\n"
    "#example config rule:\n"
    "#2 1 : 2 16\n"
    "#example meaning:\n"
    "#  if number of columns of A is 2 and number of column stops is 1,\n"
    "#  then choose the const-K SGEMM kernel with 16 row stops\n";

const char kSgemvHeaderV1[] =
    "#SGEMV smart kernel config file v1.0\n"
    "#SGEMV: Single-precision general matrix-vector multiplication\n"
    "#formula: y = alpha * A * x + beta * y\n"
    "#where:   A is M x N\n"
    "#         x is N x 1\n"
    "#         y is M x 1\n"
    "#         alpha and beta are scalars\n"
    "#format: N : kernelType blockStops rowStops\n"
    "#where:\n"
    "#  N          : number of columns of A\n"
    "#  kernelType : {basic = 0, var-K = 1, const-K = 2}\n"
    "#  blockStops : (# threads per thread block) / 16\n"
    "#  rowStops   : (#rows of A assigned to a thread block) / (blockStops*16)\n"
    "#example config rule:\n"
    "#1 : 2 3 4\n"
    "#example meaning:\n"
    "#  if number of columns of A is 1, then choose the\n"
    "#  const-K SGEMV kernel with 3 block stops and 4 row stops\n";

const char kSgemmHeaderV2[] =
    "#SGEMM smart kernel config file v2.0 (MI355X gfx950)\n"
    "#formula: C = alpha * op(A) * op(B) + beta * C, oriented so that M >= N\n"
    "#format: form log2M log2N log2K : family rt ct ksplit wpb\n"
    "#where:\n"
    "#  form   : 0 = streamed operand contiguous along K, 1 = along M\n"
    "#  log2X  : floor(log2(X)) of the oriented problem\n"
    "#  family : 1 = MFMA dot, 2 = MFMA axpy, 3 = GEMV kernels (N == 1)\n"
    "#  rt, ct : row tiles / 16-column tiles per wave\n"
    "#  ksplit : K chunks per tile, wpb : waves per workgroup (dot: they split K)\n";

const char kSgemvHeaderV2[] =
    "#SGEMV smart kernel config file v2.0 (MI355X gfx950)\n"
    "#formula: y = alpha * op(A) * x + beta * y\n"
    "#format: family log2outputs log2reduced : lpr rb split wpb\n"
    "#where:\n"
    "#  family : 1 = outputs are contiguous rows (dot), 2 = columns (axpy)\n"
    "#  lpr    : lanes per row / column chunk, rb : rows per lane group\n"
    "#  split  : split of the reduced dimension, wpb : waves per workgroup\n";

int ilog2(int64_t x) {
  int l = 0;
  while (x > 1) { x >>= 1; ++l; }
  return l;
}

bool read_lines(const char *path, std::vector<std::string> *lines) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  char buf[1024];  // cHmlLineBufferSize, include/hml_consts.h:22
  while (fgets(buf, sizeof buf, f)) lines->push_back(buf);
  fclose(f);
  return true;
}

bool is_v2(const std::vector<std::string> &lines, const char *what) {
  for (const std::string &l : lines) {
    if (l.empty() || l[0] != '#') break;
    if (l.find(what) != std::string::npos && l.find("v2.0") != std::string::npos) return true;
  }
  return false;
}

}  // namespace

// ---- parsing ------------------------------------------------------------------

hmlStatus config_load_sgemm(const char *path) {
  std::vector<std::string> lines;
  if (!read_lines(path, &lines)) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", path);
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (is_v2(lines, "SGEMM")) {
    std::vector<SgemmRuleV2> rules;
    for (const std::string &l : lines) {
      if (l.empty() || l[0] == '#') continue;
      SgemmRuleV2 r;
      if (sscanf(l.c_str(), "%d %d %d %d : %d %d %d %d %d", &r.form, &r.mlog,
                 &r.nlog, &r.klog, &r.v.family, &r.v.rt, &r.v.ct, &r.v.ksplit,
                 &r.v.wpb) != 9)
        continue;
      if (r.form < 0 || r.form > 1) continue;
      rules.push_back(r);
    }
    g_sgemm_v2 = rules;
    g_sgemm_v1_loaded = false;
  } else {
    memset(g_sgemm_v1, 0, sizeof g_sgemm_v1);
    for (const std::string &l : lines) {
      if (l.empty() || l[0] == '#') continue;
      int K, cs, type, rs;
      if (sscanf(l.c_str(), "%d %d : %d %d", &K, &cs, &type, &rs) != 4) continue;
      if (K < 0 || K > kV1MaxSkinnyK || cs < 0 || cs > kV1MaxStops) continue;
      g_sgemm_v1[K][cs].type = type;
      g_sgemm_v1[K][cs].rowStops = rs;
    }
    g_sgemm_v1_loaded = true;
    g_sgemm_v2.clear();
  }
  g_sgemm_config_done = true;
  return HML_STATUS_SUCCESS;
}

hmlStatus config_load_sgemv(const char *path) {
  std::vector<std::string> lines;
  if (!read_lines(path, &lines)) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", path);
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (is_v2(lines, "SGEMV")) {
    std::vector<SgemvRuleV2> rules;
    for (const std::string &l : lines) {
      if (l.empty() || l[0] == '#') continue;
      SgemvRuleV2 r;
      if (sscanf(l.c_str(), "%d %d %d : %d %d %d %d", &r.family, &r.olog,
                 &r.rlog, &r.v.lpr, &r.v.rb, &r.v.split, &r.v.wpb) != 7)
        continue;
      if (r.family < 1 || r.family > 2) continue;
      r.v.family = r.family;
      rules.push_back(r);
    }
    g_sgemv_v2 = rules;
    g_sgemv_v1_loaded = false;
  } else {
    memset(g_sgemv_v1, 0, sizeof g_sgemv_v1);
    for (const std::string &l : lines) {
      if (l.empty() || l[0] == '#') continue;
      int N, type, bs, rs;
      if (sscanf(l.c_str(), "%d : %d %d %d", &N, &type, &bs, &rs) != 4) continue;
      if (N < 0 || N > kV1MaxSkinnyN) continue;
      g_sgemv_v1[N].type = type;
      g_sgemv_v1[N].blockStops = bs;
      g_sgemv_v1[N].rowStops = rs;
    }
    g_sgemv_v1_loaded = true;
    g_sgemv_v2.clear();
  }
  g_sgemv_config_done = true;
  return HML_STATUS_SUCCESS;
}

hmlStatus config_save_sgemm(const char *path) {
  FILE *f = fopen(path, "wb");
  if (!f) return fail(HML_STATUS_IO_ERROR, "cannot create '%s'", path);
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (!g_sgemm_v2.empty() || !g_sgemm_v1_loaded) {
    fputs(kSgemmHeaderV2, f);
    for (const SgemmRuleV2 &r : g_sgemm_v2)
      fprintf(f, "%d %d %d %d : %d %d %d %d %d\n", r.form, r.mlog, r.nlog,
              r.klog, r.v.family, r.v.rt, r.v.ct, r.v.ksplit, r.v.wpb);
  } else {
    // hmlSgemmKernelConfigPrint, src/hml_sgemm_kernel.cu:557-568
    fputs(kSgemmHeaderV1, f);
    for (int K = 2; K <= kV1MaxSkinnyK; ++K)
      for (int cs = 1; cs <= kV1MaxStops; ++cs)
        if (g_sgemm_v1[K][cs].rowStops > 0)
          fprintf(f, "%d %d : %d %d\n", K, cs, g_sgemm_v1[K][cs].type,
                  g_sgemm_v1[K][cs].rowStops);
  }
  fclose(f);
  return HML_STATUS_SUCCESS;
}

hmlStatus config_save_sgemv(const char *path) {
  FILE *f = fopen(path, "wb");
  if (!f) return fail(HML_STATUS_IO_ERROR, "cannot create '%s'", path);
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (!g_sgemv_v2.empty() || !g_sgemv_v1_loaded) {
    fputs(kSgemvHeaderV2, f);
    for (const SgemvRuleV2 &r : g_sgemv_v2)
      fprintf(f, "%d %d %d : %d %d %d %d\n", r.family, r.olog, r.rlog,
              r.v.lpr, r.v.rb, r.v.split, r.v.wpb);
  } else {
    // hmlSgemvKernelConfigPrint, src/hml_sgemv_kernel.cu:634-645
    fputs(kSgemvHeaderV1, f);
    for (int N = 0; N <= kV1MaxSkinnyN; ++N)
      if (g_sgemv_v1[N].rowStops > 0)
        fprintf(f, "%d : %d %d %d\n", N, g_sgemv_v1[N].type,
                g_sgemv_v1[N].blockStops, g_sgemv_v1[N].rowStops);
  }
  fclose(f);
  return HML_STATUS_SUCCESS;
}

void config_clear() {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  memset(g_sgemm_v1, 0, sizeof g_sgemm_v1);
  memset(g_sgemv_v1, 0, sizeof g_sgemv_v1);
  g_sgemm_v1_loaded = g_sgemv_v1_loaded = false;
  g_sgemm_v2.clear();
  g_sgemv_v2.clear();
  g_sgemm_config_done = g_sgemv_config_done = false;
}

void config_set_sgemm_v2(const std::vector<SgemmRuleV2> &rules) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  g_sgemm_v2 = rules;
  g_sgemm_v1_loaded = false;
  g_sgemm_config_done = true;
}

void config_set_sgemv_v2(const std::vector<SgemvRuleV2> &rules) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  g_sgemv_v2 = rules;
  g_sgemv_v1_loaded = false;
  g_sgemv_config_done = true;
}

void config_add_sgemm_v2(const SgemmRuleV2 &r) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  for (SgemmRuleV2 &q : g_sgemm_v2)
    if (q.form == r.form && q.mlog == r.mlog && q.nlog == r.nlog && q.klog == r.klog &&
        (q.v.family == HML_SGEMM_FAMILY_GEMV) == (r.v.family == HML_SGEMM_FAMILY_GEMV)) {
      q = r;
      return;
    }
  g_sgemm_v2.push_back(r);
  g_sgemm_v1_loaded = false;
  g_sgemm_config_done = true;
}

void config_add_sgemv_v2(const SgemvRuleV2 &r) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  for (SgemvRuleV2 &q : g_sgemv_v2)
    if (q.family == r.family && q.olog == r.olog && q.rlog == r.rlog) {
      q = r;
      return;
    }
  g_sgemv_v2.push_back(r);
  g_sgemv_v1_loaded = false;
  g_sgemv_config_done = true;
}

bool sgemm_config_done() { return g_sgemm_config_done; }
bool sgemv_config_done() { return g_sgemv_config_done; }

int config_sgemm_v1_get(int K, int colStops, int *type, int *rowStops) {
  if (K < 0 || K > kV1MaxSkinnyK || colStops < 0 || colStops > kV1MaxStops) return 0;
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  *type = g_sgemm_v1[K][colStops].type;
  *rowStops = g_sgemm_v1[K][colStops].rowStops;
  return g_sgemm_v1_loaded ? 1 : 0;
}

int config_sgemv_v1_get(int N, int *type, int *blockStops, int *rowStops) {
  if (N < 0 || N > kV1MaxSkinnyN) return 0;
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  *type = g_sgemv_v1[N].type;
  *blockStops = g_sgemv_v1[N].blockStops;
  *rowStops = g_sgemv_v1[N].rowStops;
  return g_sgemv_v1_loaded ? 1 : 0;
}

size_t config_sgemm_v2_count() { return g_sgemm_v2.size(); }
size_t config_sgemv_v2_count() { return g_sgemv_v2.size(); }

// ---- selection from tables ---------------------------------------------------

// v1 SGEMM rule -> MI355X variant.  The key is computed as the reference
// does (src/hml_sgemm_kernel.cu:197-211): colStops = ceil(min(N,96)/16),
// K clamped to 96.  A valid rule keeps its kernel family (var-K/const-K map
// to the MFMA kernel of the layout's form, basic to the basic variant);
// rowStops sets the row tiles per wave (rowStops 1-3 -> 1, 4-7 -> 2,
// >= 8 -> 4 for the dot form), colStops the column tiles.
bool v1_sgemm_variant(const GemmKey &k, hmlSgemmVariant *v) {
  const int cs = (int)ceil_div(std::min<int64_t>(k.N, kV1MaxSkinnyN_Gemm), 16);
  const int K = (int)std::min<int64_t>(k.K, kV1MaxSkinnyK);
  const SgemmRuleV1 r = g_sgemm_v1[K][cs];
  if (r.rowStops <= 0 || r.type == 0) {
    basic_sgemm(k, v);
    return true;
  }
  heuristic_sgemm(k, v);
  if (v->family == HML_SGEMM_FAMILY_GEMV || v->family == HML_SGEMM_FAMILY_GEMVN) return true;
  const int ctcap = k.form == kFormDot ? 4 : 8;
  int ct = 1;
  while (ct < ctcap && ct < cs) ct <<= 1;
  v->ct = ct;
  if (k.form == kFormDot) {
    v->rt = r.rowStops >= 8 ? 4 : r.rowStops >= 4 ? 2 : 1;
    while (v->rt * v->ct > 8) v->rt >>= 1;
    if (!skinny_dot_runnable(v->rt, v->ct, v->wpb)) v->wpb = 4;
  } else {
    v->rt = r.rowStops >= 8 ? 2 : 1;
  }
  return true;
}

bool config_select_sgemm(const GemmKey &k, hmlSgemmVariant *v, int *source, bool exact_only) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (!g_sgemm_v2.empty()) {
    const int ml = ilog2(k.M), nl = ilog2(k.N), kl = ilog2(k.K);
    int best = -1, bestd = 1 << 30;
    for (size_t i = 0; i < g_sgemm_v2.size(); ++i) {
      const SgemmRuleV2 &r = g_sgemm_v2[i];
      if (r.form != k.form) continue;
      if ((r.v.family == HML_SGEMM_FAMILY_GEMV) != k.gemv_ok) continue;
      const int d = abs(r.mlog - ml) + 2 * abs(r.nlog - nl) + abs(r.klog - kl);
      if (d < bestd) { bestd = d; best = (int)i; }
    }
    if (best >= 0 && (!exact_only || bestd == 0)) {
      *v = g_sgemm_v2[best].v;
      *source = 1;
      return true;
    }
    return false;
  }
  if (g_sgemm_v1_loaded && k.order == HML_ROW_MAJOR && k.transA == HML_OP_N &&
      k.transB == HML_OP_N) {
    *source = 1;
    return v1_sgemm_variant(k, v);
  }
  return false;
}

bool config_select_sgemv(const GemvKey &k, hmlSgemvVariant *v, int *source, bool exact_only) {
  std::lock_guard<std::mutex> lk(g_cfg_mutex);
  if (!g_sgemv_v2.empty()) {
    const int ol = ilog2(k.outputs), rl = ilog2(k.reduced);
    int best = -1, bestd = 1 << 30;
    for (size_t i = 0; i < g_sgemv_v2.size(); ++i) {
      const SgemvRuleV2 &r = g_sgemv_v2[i];
      if (r.family != k.family) continue;
      const int d = abs(r.olog - ol) + abs(r.rlog - rl);
      if (d < bestd) { bestd = d; best = (int)i; }
    }
    if (best >= 0 && (!exact_only || bestd == 0)) {
      *v = g_sgemv_v2[best].v;
      *source = 1;
      return true;
    }
    return false;
  }
  if (g_sgemv_v1_loaded && k.family == HML_SGEMV_FAMILY_DOT) {
    // rule = cfg[N <= 1024 ? N : 0] (src/hml_sgemv_kernel.cu:322-330)
    const int N = k.reduced <= kV1MaxSkinnyN ? (int)k.reduced : 0;
    const SgemvRuleV1 r = g_sgemv_v1[N];
    *source = 1;
    if (r.blockStops <= 0 || r.rowStops <= 0 || r.type == 0) {
      basic_sgemv(k, v);
      return true;
    }
    heuristic_sgemv(k, v);
    int rb = 1;
    while (rb < 8 && rb * 2 <= r.rowStops) rb <<= 1;
    v->rb = rb;
    return true;
  }
  return false;
}

}  // namespace hml
