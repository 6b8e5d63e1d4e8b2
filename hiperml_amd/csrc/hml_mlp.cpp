// hml_mlp.cpp -- the learned kernel selector: an FP32 MLP trained on the
// GPU by the kernels of hml_k_mlp.hip, its class space (class <-> kernel
// variant), and its use by the SGEMM dispatch.
//
// Replaces the static table lookup of the reference (hmlSgemmKernelSelect,
// src/hml_sgemm_kernel.cu:189-237) as the first selector when a trained net
// is installed; the config table and the built-in heuristic stay as
// fallbacks (a pick that cannot run the shape is never used).  Decisions
// are memoised per problem key, so the network runs once per new shape.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "hml_internal.h"
#include "hml_mlp.h"
#include "hml_runtime.h"
#include "hml_select.h"

struct hmlMlp_t {
  int n_in = 0, hidden = 0, n_hidden = 0, n_out = 0;
  std::vector<int> dims;         // n_in, hidden..., n_out
  std::vector<int64_t> w_off;    // offset of W_l in the parameter vector
  std::vector<int64_t> b_off;    // offset of b_l
  int64_t nparam = 0;
  float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
  int64_t adam_t = 0;
  // activations / gradients for the largest batch seen
  int cap = 0;
  std::vector<float *> act;      // act[0] = input, act[l] = output of layer l
  std::vector<float *> dz;       // dz[l] = dLoss/d(pre-activation of layer l)
  float *targets = nullptr, *mask = nullptr, *probs = nullptr, *loss_row = nullptr, *loss = nullptr;
  hipStream_t st = nullptr;
};

namespace hml {
namespace {

uint64_t splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double u01(uint64_t *s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

void free_batch(hmlMlp_t *n) {
  for (size_t l = 0; l < n->act.size(); ++l) if (n->act[l]) (void)hipFree(n->act[l]);
  for (size_t l = 0; l < n->dz.size(); ++l) if (n->dz[l]) (void)hipFree(n->dz[l]);
  n->act.assign(n->dims.size(), nullptr);
  n->dz.assign(n->dims.size(), nullptr);
  for (float **p : {&n->targets, &n->mask, &n->probs, &n->loss_row, &n->loss})
    if (*p) { (void)hipFree(*p); *p = nullptr; }
  n->cap = 0;
}

hmlStatus ensure_batch(hmlMlp_t *n, int B) {
  if (B <= n->cap) return HML_STATUS_SUCCESS;
  free_batch(n);
  const int L = (int)n->dims.size() - 1;
  hipError_t e = hipSuccess;
  for (int l = 0; l <= L && e == hipSuccess; ++l) {
    e = hipMalloc(&n->act[l], sizeof(float) * (size_t)B * n->dims[l]);
    if (e == hipSuccess && l > 0) e = hipMalloc(&n->dz[l], sizeof(float) * (size_t)B * n->dims[l]);
  }
  const size_t out = sizeof(float) * (size_t)B * n->n_out;
  if (e == hipSuccess) e = hipMalloc(&n->targets, out);
  if (e == hipSuccess) e = hipMalloc(&n->mask, out);
  if (e == hipSuccess) e = hipMalloc(&n->probs, out);
  if (e == hipSuccess) e = hipMalloc(&n->loss_row, sizeof(float) * B);
  if (e == hipSuccess) e = hipMalloc(&n->loss, sizeof(float));
  if (e != hipSuccess) {
    free_batch(n);
    return fail(HML_STATUS_ALLOC_FAILED, "mlp batch buffers: %s", hipGetErrorString(e));
  }
  n->cap = B;
  return HML_STATUS_SUCCESS;
}

// Forward pass of act[0] (B rows); logits land in act[L].
hipError_t forward(hmlMlp_t *n, int B) {
  const int L = (int)n->dims.size() - 1;
  for (int l = 1; l <= L; ++l) {
    hipError_t e = mlp_linear_fwd(n->act[l - 1], n->params + n->w_off[l], n->params + n->b_off[l],
                                  n->act[l], B, n->dims[l - 1], n->dims[l], l < L ? 1 : 0, n->st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Loss and gradients of the batch in act[0] / targets / mask.
hipError_t backward(hmlMlp_t *n, int B) {
  const int L = (int)n->dims.size() - 1;
  hipError_t e = mlp_softmax_xent(n->act[L], n->targets, n->mask, n->probs, n->dz[L],
                                  n->loss_row, B, n->n_out, n->st);
  for (int l = L; l >= 1 && e == hipSuccess; --l) {
    e = mlp_linear_bwd_weight(n->act[l - 1], n->dz[l], n->grads + n->w_off[l],
                              n->grads + n->b_off[l], B, n->dims[l - 1], n->dims[l], n->st);
    if (e == hipSuccess && l > 1)
      e = mlp_linear_bwd_data(n->dz[l], n->params + n->w_off[l], n->act[l - 1], n->dz[l - 1],
                              B, n->dims[l - 1], n->dims[l], n->st);
  }
  if (e == hipSuccess) e = mlp_sum(n->loss_row, B, n->loss, n->st);
  return e;
}

hipError_t adam_step(hmlMlp_t *n, float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->adam_t += 1;
  const double t = (double)n->adam_t;
  const float lr_t = (float)(lr * sqrt(1.0 - pow(b2, t)) / (1.0 - pow(b1, t)));
  return mlp_adam(n->params, n->grads, n->adam_m, n->adam_v, (int)n->nparam, lr_t, b1, b2, eps, n->st);
}

// ---- selector ------------------------------------------------------------------

// Host copy of the installed selector network.  Selection runs its forward
// pass on the host (about 13k multiply-adds per new problem), so a GEMM
// call never allocates, synchronises or launches anything for it.
struct HostNet {
  std::vector<int> dims;      // n_in, hidden..., n_out
  std::vector<float> params;  // per layer W[in][out] then bias[out]
};

std::mutex g_sel_mutex;
std::shared_ptr<const HostNet> g_sel_host;  // the installed selector
hmlMlp_t *g_sel_net = nullptr;              // device net it was copied from
std::unordered_map<std::string, hmlSgemmVariant> g_sel_memo;
// Decisions memoised per problem key; a process that sees more distinct
// problems than this starts the memo afresh.
constexpr size_t kSelMemoCap = 4096;

std::shared_ptr<const HostNet> host_copy(const std::vector<int> &dims,
                                         std::vector<float> params) {
  auto h = std::make_shared<HostNet>();
  h->dims = dims;
  h->params = std::move(params);
  return h;
}

// Logits of one input row: ReLU after every layer but the last, sums in
// double in input order.
void host_forward(const HostNet &n, const float *x, double *logits) {
  std::vector<double> a(x, x + n.dims[0]), z;
  size_t off = 0;
  const int L = (int)n.dims.size() - 1;
  for (int l = 1; l <= L; ++l) {
    const int I = n.dims[l - 1], O = n.dims[l];
    const float *W = n.params.data() + off;
    const float *b = W + (size_t)I * O;
    z.assign(O, 0.0);
    for (int o = 0; o < O; ++o) {
      double acc = b[o];
      for (int i = 0; i < I; ++i) acc += a[i] * (double)W[(size_t)i * O + o];
      z[o] = (l < L && acc < 0.0) ? 0.0 : acc;
    }
    off += (size_t)I * O + O;
    a.swap(z);
  }
  for (size_t o = 0; o < a.size(); ++o) logits[o] = a[o];
}

// Reads an MLP file: header {magic "HMLP", n_in, hidden, n_hidden, n_out},
// then the parameters as floats.
hmlStatus read_mlp_file(const char *path, int32_t hdr[5], std::vector<float> *params) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail(HML_STATUS_IO_ERROR, "cannot open '%s'", path);
  if (fread(hdr, sizeof(int32_t) * 5, 1, f) != 1 || hdr[0] != 0x504c4d48 || hdr[1] <= 0 ||
      hdr[2] <= 0 || hdr[3] < 0 || hdr[3] > 64 || hdr[4] <= 1 || hdr[1] > 4096 ||
      hdr[2] > 4096 || hdr[4] > 4096) {
    fclose(f);
    return fail(HML_STATUS_PARSE_ERROR, "'%s' is not an MLP file", path);
  }
  int64_t count = 0, in = hdr[1];
  for (int l = 0; l <= hdr[3]; ++l) {
    const int64_t out = l < hdr[3] ? hdr[2] : hdr[4];
    count += in * out + out;
    in = out;
  }
  params->resize((size_t)count);
  const size_t got = fread(params->data(), sizeof(float), params->size(), f);
  fclose(f);
  if (got != params->size()) return fail(HML_STATUS_PARSE_ERROR, "'%s' is truncated", path);
  return HML_STATUS_SUCCESS;
}

// Re-reads the parameters of the installed device net (after training or a
// parameter update).  Caller holds g_sel_mutex.
void refresh_host_copy_locked(hmlMlp_t *n) {
  std::vector<float> h((size_t)n->nparam);
  if (hipStreamSynchronize(n->st) != hipSuccess ||
      hipMemcpy(h.data(), n->params, sizeof(float) * h.size(), hipMemcpyDeviceToHost) !=
          hipSuccess) {
    g_sel_host.reset();
    g_sel_net = nullptr;
  } else {
    g_sel_host = host_copy(n->dims, std::move(h));
  }
  g_sel_memo.clear();
}

int ctmax_for(int64_t N) {
  const int64_t nct = (N + 15) / 16;
  int c = 1;
  while (c < 8 && c < nct) c <<= 1;
  return c;
}

}  // namespace

// Row-major equivalent of the caller's problem: a column-major C = op(A)op(B)
// is the row-major C^T = op(B)^T op(A)^T.
void selector_features(const GemmKey &k, float *f) {
  int64_t M = k.cM, N = k.cN;
  int ta = k.transA, tb = k.transB;
  if (k.order == HML_COL_MAJOR) {
    // The first row-major operand is op(B)^T, stored exactly as B is
    // (a K x N column-major B is an N x K row-major B^T), so the transpose
    // flags swap along with M and N.
    std::swap(M, N);
    std::swap(ta, tb);
  }
  f[0] = (float)(log2((double)std::max<int64_t>(M, 1)) / 16.0);
  f[1] = (float)(log2((double)std::max<int64_t>(N, 1)) / 16.0);
  f[2] = (float)(log2((double)std::max<int64_t>(k.K, 1)) / 16.0);
  f[3] = (float)ta;
  f[4] = (float)tb;
  f[5] = (float)k.beta0;
}

// Class -> concrete variant for this key; false when it cannot run it.
bool selector_variant(const GemmKey &k, int cls, hmlSgemmVariant *v) {
  if (cls < 0 || cls >= kSelClasses) return false;
  if (cls == 84) {
    *v = hmlSgemmVariant{HML_SGEMM_FAMILY_DENSE, 1, 1, 1, 4};
    return k.N >= 32 && k.M >= 256 && k.K >= 64 && sgemm_variant_runnable(k, *v);
  }
  if (cls == 85) {
    if (!k.gemv_ok) return false;
    GemmKey g = k;
    heuristic_sgemm(g, v);
    return v->family == HML_SGEMM_FAMILY_GEMV && sgemm_variant_runnable(k, *v);
  }
  if (cls >= kSelGemvDotFirst) {
    // explicit GEMV geometries (N == 1 with unit-stride x and y)
    if (!k.gemv_ok) return false;
    const bool dot = cls < kSelGemvAxpyFirst;
    if (dot != (k.form == kFormDot)) return false;
    int lpr, rb, wpb;
    if (dot) {
      const int c = cls - kSelGemvDotFirst;  // lpr x rb x wpb = 4 x 3 x 2
      lpr = 8 << (c / 6);
      rb = 1 << ((c / 2) % 3);
      wpb = (c & 1) ? 8 : 2;
    } else {
      const int c = cls - kSelGemvAxpyFirst;  // lpr x wpb = 4 x 2
      lpr = 8 << (c / 2);
      rb = 1;
      wpb = (c & 1) ? 8 : 4;
    }
    const GemvKey gk{dot ? HML_SGEMV_FAMILY_DOT : HML_SGEMV_FAMILY_AXPY, k.M, k.K, k.beta0};
    *v = hmlSgemmVariant{HML_SGEMM_FAMILY_GEMV, lpr, rb, gemv_split_rule(gk, lpr, rb), wpb};
    return sgemm_variant_runnable(k, *v);
  }
  if (cls == 86) {
    if (k.form != kFormDot || !(k.N == 2 || k.N == 4 || k.N == 8)) return false;
    GemvKey gk{HML_SGEMV_FAMILY_DOT, k.M, k.K, k.beta0};
    hmlSgemvVariant gv;
    heuristic_sgemv(gk, &gv);
    *v = hmlSgemmVariant{HML_SGEMM_FAMILY_GEMVN, gv.lpr, std::min(gv.rb, 4), gv.split, gv.wpb};
    return sgemm_variant_runnable(k, *v);
  }
  const int wpb_i = cls & 1;
  const int ks_i = (cls >> 1) % 7;
  const int rest = (cls >> 1) / 7;  // rt_i * 2 + ctm
  const int ctm = rest & 1, rt_i = rest >> 1;
  // DOT tiles cover at most 4 column tiles (more columns take more
  // workgroups); AXPY tiles up to 8
  const int ctmax = std::min(ctmax_for(k.N), k.form == kFormDot ? 4 : 8);
  if (ctm == 1 && ctmax == 1) return false;
  v->family = k.form == kFormDot ? HML_SGEMM_FAMILY_DOT : HML_SGEMM_FAMILY_AXPY;
  v->rt = 1 << rt_i;
  v->ct = ctm ? ctmax / 2 : ctmax;
  v->ksplit = 1 << ks_i;
  v->wpb = wpb_i ? 8 : 4;
  if (v->rt * v->ct > 16) return false;
  const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(64, k.K / 128));
  if (v->ksplit > cap) return false;
  return sgemm_variant_runnable(k, *v);
}

int selector_class_of(const GemmKey &k, const hmlSgemmVariant &v) {
  for (int c = 0; c < kSelClasses; ++c) {
    hmlSgemmVariant u;
    if (selector_variant(k, c, &u) && u.family == v.family && u.rt == v.rt &&
        u.ct == v.ct && u.ksplit == v.ksplit && u.wpb == v.wpb)
      return c;
  }
  return -1;
}

// The shipped selector: $HML_SELECTOR (a model file, or "off"), else
// config/hml_selector_mi355x.bin beside the hiperml_amd/ directory that
// holds this library.  Loaded once per process, on the first selection (on
// the device current then), and installed unless the caller installed a
// network first; a missing or unreadable file leaves the table / heuristic
// selection in place.
std::once_flag g_autoload_once;

void selector_autoload() {
  std::call_once(g_autoload_once, [] {
    // On by default: unset, empty or "shipped" loads the model shipped
    // beside the library (../config/hml_selector_mi355x.bin); "off" or "0"
    // keeps the table and the heuristic; any other value names a model
    // file.
    const char *env = getenv("HML_SELECTOR");
    if (env && (!strcmp(env, "off") || !strcmp(env, "0"))) return;
    std::string path;
    if (env && *env && strcmp(env, "shipped") != 0) {
      path = env;
    } else {
      Dl_info info;
      if (!dladdr((void *)&hmlMlpLoad, &info) || !info.dli_fname) return;
      std::string lib = info.dli_fname;
      const size_t sl = lib.rfind('/');
      path = (sl == std::string::npos ? std::string(".") : lib.substr(0, sl)) +
             "/../config/hml_selector_mi355x.bin";
    }
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return;
    fclose(f);
    int32_t hdr[5];
    std::vector<float> params;
    if (read_mlp_file(path.c_str(), hdr, &params) != HML_STATUS_SUCCESS) {
      if (verbosity() >= 1)
        fprintf(stderr, "Warning: selector '%s' not loaded: %s\n", path.c_str(), hmlLastError());
      clear_error();
      return;
    }
    if (hdr[1] != kSelFeatures || hdr[4] != kSelClasses) {
      if (verbosity() >= 1)
        fprintf(stderr, "Warning: selector '%s' maps %d features to %d classes, not %d to %d\n",
                path.c_str(), hdr[1], hdr[4], kSelFeatures, kSelClasses);
      return;
    }
    std::vector<int> dims{hdr[1]};
    for (int i = 0; i < hdr[3]; ++i) dims.push_back(hdr[2]);
    dims.push_back(hdr[4]);
    std::lock_guard<std::mutex> lk(g_sel_mutex);
    if (!g_sel_host) {  // a net installed by the caller first is kept
      g_sel_host = host_copy(dims, std::move(params));
      g_sel_memo.clear();
    }
  });
}

bool learned_select_sgemm(const GemmKey &k, hmlSgemmVariant *v, bool gemv_only) {
  selector_autoload();
  std::shared_ptr<const HostNet> net;
  char key[168];
  snprintf(key, sizeof key, "%d:%lld:%lld:%lld:%d:%d:%d:%d:%d:%lld:%lld:%d", k.form,
           (long long)k.M, (long long)k.N, (long long)k.K, k.beta0, k.transA,
           k.transB, k.order, (int)k.gemv_ok, (long long)k.cM, (long long)k.cN,
           (int)gemv_only);
  {
    std::lock_guard<std::mutex> lk(g_sel_mutex);
    if (!g_sel_host) return false;
    auto it = g_sel_memo.find(key);
    if (it != g_sel_memo.end()) {
      *v = it->second;
      return true;
    }
    net = g_sel_host;
  }
  float x[kSelFeatures];
  double logits[kSelClasses];
  selector_features(k, x);
  host_forward(*net, x, logits);
  // argmax of the masked softmax = argmax of the allowed logits
  int best = -1;
  hmlSgemmVariant pick{};
  for (int c = 0; c < kSelClasses; ++c) {
    hmlSgemmVariant u;
    if (!selector_variant(k, c, &u)) continue;
    if (gemv_only && u.family != HML_SGEMM_FAMILY_GEMV) continue;
    if (best < 0 || logits[c] > logits[best]) {
      best = c;
      pick = u;
    }
  }
  if (best < 0) return false;
  *v = pick;
  std::lock_guard<std::mutex> lk(g_sel_mutex);
  if (g_sel_host == net) {
    if (g_sel_memo.size() >= kSelMemoCap) g_sel_memo.clear();
    g_sel_memo[key] = pick;
  }
  return true;
}

// SGEMV through the selector: the problem as the GEMM it is (N = 1, x as a
// unit-stride column; row-major A x is the dot form, A^T x the axpy form),
// choosing among the GEMV kernel classes.
bool learned_select_sgemv(const GemvKey &gk, hmlSgemvVariant *v) {
  GemmKey k;
  k.form = gk.family == HML_SGEMV_FAMILY_DOT ? kFormDot : kFormAxpy;
  k.M = gk.outputs; k.N = 1; k.K = gk.reduced;
  k.beta0 = gk.beta0;
  k.transA = gk.family == HML_SGEMV_FAMILY_DOT ? HML_OP_N : HML_OP_T;
  k.transB = HML_OP_N;
  k.order = HML_ROW_MAJOR;
  k.gemv_ok = true;
  k.cM = gk.outputs; k.cN = 1;
  hmlSgemmVariant u;
  if (!learned_select_sgemm(k, &u, true)) return false;
  *v = hmlSgemvVariant{gk.family, u.rt, u.ct, u.ksplit, u.wpb};
  return true;
}

}  // namespace hml

using namespace hml;

extern "C" {

hmlStatus hmlMlpCreate(hmlMlp_t **net, int32_t n_in, int32_t hidden,
                       int32_t n_hidden, int32_t n_out, uint64_t seed) {
  clear_error();
  if (!net || n_in <= 0 || hidden <= 0 || n_hidden < 0 || n_out <= 1)
    return fail(HML_STATUS_INVALID_VALUE, "mlp: bad shape");
  hmlMlp_t *n = new hmlMlp_t();
  n->n_in = n_in; n->hidden = hidden; n->n_hidden = n_hidden; n->n_out = n_out;
  n->dims.push_back(n_in);
  for (int i = 0; i < n_hidden; ++i) n->dims.push_back(hidden);
  n->dims.push_back(n_out);
  const int L = (int)n->dims.size() - 1;
  n->w_off.assign(L + 1, 0);
  n->b_off.assign(L + 1, 0);
  int64_t off = 0;
  for (int l = 1; l <= L; ++l) {
    n->w_off[l] = off; off += (int64_t)n->dims[l - 1] * n->dims[l];
    n->b_off[l] = off; off += n->dims[l];
  }
  n->nparam = off;
  // He-uniform weights, zero biases, from a seeded host stream
  std::vector<float> h(off, 0.f);
  uint64_t s = seed;
  for (int l = 1; l <= L; ++l) {
    const double lim = sqrt(6.0 / n->dims[l - 1]);
    for (int64_t i = 0; i < (int64_t)n->dims[l - 1] * n->dims[l]; ++i)
      h[n->w_off[l] + i] = (float)((2.0 * u01(&s) - 1.0) * lim);
  }
  hipError_t e = hipStreamCreate(&n->st);
  for (float **p : {&n->params, &n->grads, &n->adam_m, &n->adam_v})
    if (e == hipSuccess) e = hipMalloc(p, sizeof(float) * off);
  if (e == hipSuccess) e = hipMemcpy(n->params, h.data(), sizeof(float) * off, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(n->adam_m, 0, sizeof(float) * off);
  if (e == hipSuccess) e = hipMemset(n->adam_v, 0, sizeof(float) * off);
  if (e != hipSuccess) {
    hmlMlpDestroy(n);
    return fail(HML_STATUS_HIP_ERROR, "mlp create: %s", hipGetErrorString(e));
  }
  n->act.assign(n->dims.size(), nullptr);
  n->dz.assign(n->dims.size(), nullptr);
  *net = n;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlMlpDestroy(hmlMlp_t *n) {
  if (!n) return HML_STATUS_SUCCESS;
  {
    // destroying the installed net uninstalls it
    std::lock_guard<std::mutex> lk(g_sel_mutex);
    if (g_sel_net == n) {
      g_sel_net = nullptr;
      g_sel_host.reset();
      g_sel_memo.clear();
    }
  }
  if (n->st) (void)hipStreamSynchronize(n->st);
  free_batch(n);
  for (float *p : {n->params, n->grads, n->adam_m, n->adam_v}) if (p) (void)hipFree(p);
  if (n->st) (void)hipStreamDestroy(n->st);
  delete n;
  return HML_STATUS_SUCCESS;
}

int64_t hmlMlpParamCount(hmlMlp_t *n) { return n ? n->nparam : -1; }

hmlStatus hmlMlpGetParams(hmlMlp_t *n, float *host, int64_t count) {
  if (!n || !host || count != n->nparam) return fail(HML_STATUS_INVALID_VALUE, "mlp params");
  hipError_t e = hipStreamSynchronize(n->st);
  if (e != hipSuccess) return from_hip(e, "mlp get");
  return from_hip(hipMemcpy(host, n->params, sizeof(float) * count, hipMemcpyDeviceToHost), "mlp get");
}

hmlStatus hmlMlpSetParams(hmlMlp_t *n, const float *host, int64_t count) {
  if (!n || !host || count != n->nparam) return fail(HML_STATUS_INVALID_VALUE, "mlp params");
  hipError_t e = hipStreamSynchronize(n->st);
  if (e != hipSuccess) return from_hip(e, "mlp set");
  e = hipMemcpy(n->params, host, sizeof(float) * count, hipMemcpyHostToDevice);
  if (e != hipSuccess) return from_hip(e, "mlp set");
  std::lock_guard<std::mutex> lk(g_sel_mutex);
  if (g_sel_net == n) refresh_host_copy_locked(n);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlMlpLossGrad(hmlMlp_t *n, const float *X, const float *T,
                         const float *mask, int32_t batch, float *loss,
                         float *grads) {
  clear_error();
  if (!n || !X || !T || !mask || batch <= 0) return fail(HML_STATUS_INVALID_VALUE, "mlp lossgrad");
  hmlStatus s = ensure_batch(n, batch);
  if (s) return s;
  const size_t out = sizeof(float) * (size_t)batch * n->n_out;
  hipError_t e = hipMemcpyAsync(n->act[0], X, sizeof(float) * (size_t)batch * n->n_in, hipMemcpyHostToDevice, n->st);
  if (e == hipSuccess) e = hipMemcpyAsync(n->targets, T, out, hipMemcpyHostToDevice, n->st);
  if (e == hipSuccess) e = hipMemcpyAsync(n->mask, mask, out, hipMemcpyHostToDevice, n->st);
  if (e == hipSuccess) e = forward(n, batch);
  if (e == hipSuccess) e = backward(n, batch);
  if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->st);
  if (e == hipSuccess && grads)
    e = hipMemcpyAsync(grads, n->grads, sizeof(float) * n->nparam, hipMemcpyDeviceToHost, n->st);
  if (e == hipSuccess) e = hipStreamSynchronize(n->st);
  if (e == hipSuccess && loss) *loss /= (float)batch;
  return from_hip(e, "mlp loss/grad");
}

hmlStatus hmlMlpTrain(hmlMlp_t *n, const float *X, const float *T,
                      const float *mask, int64_t count, int32_t epochs,
                      int32_t batch, float lr, uint64_t seed,
                      float *epoch_loss) {
  clear_error();
  if (!n || !X || !T || !mask || count <= 0 || epochs <= 0 || batch <= 0 || count > 0x7fffffff)
    return fail(HML_STATUS_INVALID_VALUE, "mlp train: bad arguments");
  const int B = (int)std::min<int64_t>(batch, count);
  hmlStatus s = ensure_batch(n, B);
  if (s) return s;
  float *dX = nullptr, *dT = nullptr, *dM = nullptr, *dL = nullptr;
  int32_t *dIdx = nullptr;
  const int64_t nb = (count + B - 1) / B;
  hipError_t e = hipMalloc(&dX, sizeof(float) * count * n->n_in);
  if (e == hipSuccess) e = hipMalloc(&dT, sizeof(float) * count * n->n_out);
  if (e == hipSuccess) e = hipMalloc(&dM, sizeof(float) * count * n->n_out);
  if (e == hipSuccess) e = hipMalloc(&dIdx, sizeof(int32_t) * count);
  if (e == hipSuccess) e = hipMalloc(&dL, sizeof(float) * nb);
  if (e == hipSuccess) e = hipMemcpy(dX, X, sizeof(float) * count * n->n_in, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dT, T, sizeof(float) * count * n->n_out, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dM, mask, sizeof(float) * count * n->n_out, hipMemcpyHostToDevice);
  std::vector<int32_t> idx(count);
  std::iota(idx.begin(), idx.end(), 0);
  std::vector<float> bl(nb);
  uint64_t rs = seed;
  for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    for (int64_t i = count - 1; i > 0; --i) {  // Fisher-Yates
      const int64_t j = (int64_t)(splitmix(&rs) % (uint64_t)(i + 1));
      std::swap(idx[i], idx[j]);
    }
    e = hipMemcpyAsync(dIdx, idx.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, n->st);
    for (int64_t b = 0; b < nb && e == hipSuccess; ++b) {
      const int rows = (int)std::min<int64_t>(B, count - b * B);
      e = mlp_gather_rows(dX, dIdx + b * B, n->act[0], rows, n->n_in, n->st);
      if (e == hipSuccess) e = mlp_gather_rows(dT, dIdx + b * B, n->targets, rows, n->n_out, n->st);
      if (e == hipSuccess) e = mlp_gather_rows(dM, dIdx + b * B, n->mask, rows, n->n_out, n->st);
      if (e == hipSuccess) e = forward(n, rows);
      if (e == hipSuccess) e = backward(n, rows);
      if (e == hipSuccess) e = adam_step(n, lr);
      if (e == hipSuccess)
        e = hipMemcpyAsync(dL + b, n->loss, sizeof(float), hipMemcpyDeviceToDevice, n->st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(bl.data(), dL, sizeof(float) * nb, hipMemcpyDeviceToHost, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
    if (e == hipSuccess && epoch_loss) {
      double tot = 0.0;
      for (float v : bl) tot += v;
      epoch_loss[ep] = (float)(tot / (double)count);
    }
  }
  for (void *p : {(void *)dX, (void *)dT, (void *)dM, (void *)dIdx, (void *)dL}) if (p) (void)hipFree(p);
  {
    std::lock_guard<std::mutex> lk(g_sel_mutex);
    if (g_sel_net == n) refresh_host_copy_locked(n);
  }
  return from_hip(e, "mlp train");
}

hmlStatus hmlMlpPredict(hmlMlp_t *n, const float *X, const float *mask,
                        int64_t count, float *probs) {
  if (!n || !X || !mask || !probs || count <= 0) return fail(HML_STATUS_INVALID_VALUE, "mlp predict");
  const int B = (int)std::min<int64_t>(count, 4096);
  hmlStatus s = ensure_batch(n, B);
  if (s) return s;
  const int L = (int)n->dims.size() - 1;
  hipError_t e = hipSuccess;
  for (int64_t r0 = 0; r0 < count && e == hipSuccess; r0 += B) {
    const int rows = (int)std::min<int64_t>(B, count - r0);
    const size_t out = sizeof(float) * (size_t)rows * n->n_out;
    e = hipMemcpyAsync(n->act[0], X + r0 * n->n_in, sizeof(float) * rows * n->n_in, hipMemcpyHostToDevice, n->st);
    if (e == hipSuccess) e = hipMemcpyAsync(n->mask, mask + r0 * n->n_out, out, hipMemcpyHostToDevice, n->st);
    if (e == hipSuccess) e = forward(n, rows);
    if (e == hipSuccess)
      e = mlp_softmax_xent(n->act[L], nullptr, n->mask, n->probs, nullptr, nullptr, rows, n->n_out, n->st);
    if (e == hipSuccess) e = hipMemcpyAsync(probs + r0 * n->n_out, n->probs, out, hipMemcpyDeviceToHost, n->st);
    if (e == hipSuccess) e = hipStreamSynchronize(n->st);
  }
  return from_hip(e, "mlp predict");
}

hmlStatus hmlMlpSave(hmlMlp_t *n, const char *path) {
  if (!n || !path) return fail(HML_STATUS_INVALID_VALUE, "mlp save");
  std::vector<float> h(n->nparam);
  hmlStatus s = hmlMlpGetParams(n, h.data(), n->nparam);
  if (s) return s;
  FILE *f = fopen(path, "wb");
  if (!f) return fail(HML_STATUS_IO_ERROR, "cannot create '%s'", path);
  const int32_t hdr[5] = {0x504c4d48 /* "HMLP" */, n->n_in, n->hidden, n->n_hidden, n->n_out};
  fwrite(hdr, sizeof hdr, 1, f);
  fwrite(h.data(), sizeof(float), h.size(), f);
  fclose(f);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlMlpLoad(hmlMlp_t **net, const char *path) {
  if (!net || !path) return fail(HML_STATUS_INVALID_VALUE, "mlp load");
  int32_t hdr[5];
  std::vector<float> h;
  hmlStatus s = read_mlp_file(path, hdr, &h);
  if (s) return s;
  s = hmlMlpCreate(net, hdr[1], hdr[2], hdr[3], hdr[4], 0);
  if (s) return s;
  if ((int64_t)h.size() != (*net)->nparam) {
    hmlMlpDestroy(*net);
    *net = nullptr;
    return fail(HML_STATUS_PARSE_ERROR, "'%s' has an inconsistent header", path);
  }
  return hmlMlpSetParams(*net, h.data(), (int64_t)h.size());
}

int32_t hmlSelectorClassCount(void) { return kSelClasses; }
int32_t hmlSelectorFeatureCount(void) { return kSelFeatures; }

static GemmKey key_of(const hmlGemmShape *s) {
  GemmCanon g = dense_canon((hmlOrder)s->order, (hmlOp)s->transA, (hmlOp)s->transB, s->M, s->N, s->K);
  return make_key(g, s->beta0 ? 0.0f : 1.0f, 0, (hmlOrder)s->order, (hmlOp)s->transA, (hmlOp)s->transB);
}

hmlStatus hmlSelectorFeatures(const hmlGemmShape *shape, float *features) {
  if (!shape || !features || shape->M <= 0 || shape->N <= 0 || shape->K <= 0)
    return fail(HML_STATUS_INVALID_VALUE, "selector features");
  selector_features(key_of(shape), features);
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSelectorMask(const hmlGemmShape *shape, float *mask) {
  if (!shape || !mask || shape->M <= 0 || shape->N <= 0 || shape->K <= 0)
    return fail(HML_STATUS_INVALID_VALUE, "selector mask");
  const GemmKey k = key_of(shape);
  for (int c = 0; c < kSelClasses; ++c) {
    hmlSgemmVariant u;
    mask[c] = selector_variant(k, c, &u) ? 1.f : 0.f;
  }
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSelectorVariant(const hmlGemmShape *shape, int32_t cls,
                             hmlSgemmVariant *v) {
  if (!shape || !v) return fail(HML_STATUS_INVALID_VALUE, "selector variant");
  if (!selector_variant(key_of(shape), cls, v))
    return fail(HML_STATUS_NO_KERNEL, "class %d cannot run this shape", cls);
  return HML_STATUS_SUCCESS;
}

int32_t hmlSelectorClassOf(const hmlGemmShape *shape, const hmlSgemmVariant *v) {
  if (!shape || !v) return -1;
  return selector_class_of(key_of(shape), *v);
}

hmlStatus hmlSelectorInstall(hmlMlp_t *net) {
  if (net && (net->n_in != kSelFeatures || net->n_out != kSelClasses))
    return fail(HML_STATUS_INVALID_VALUE, "selector net must map %d features to %d classes",
                kSelFeatures, kSelClasses);
  selector_autoload();  // so a later first selection cannot replace this choice
  std::lock_guard<std::mutex> lk(g_sel_mutex);
  g_sel_net = net;
  g_sel_memo.clear();
  if (!net) {
    g_sel_host.reset();
    return HML_STATUS_SUCCESS;
  }
  refresh_host_copy_locked(net);
  if (!g_sel_host) return fail(HML_STATUS_HIP_ERROR, "selector install: cannot read the parameters");
  return HML_STATUS_SUCCESS;
}

}  // extern "C"
