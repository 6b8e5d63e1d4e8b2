// hml_grouped.cpp -- grouped (batched) skinny SGEMM: a plan holds the device
// descriptors of many independent row-major problems and, per tile class
// (by N), a longest-first tile schedule; running it launches one persistent
// kernel per non-empty class, optionally replayed from a captured hipGraph.
// No reference counterpart (the reference runs one problem per call,
// src/hml_sgemm.cu:120-211).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "hml_internal.h"
#include "hml_runtime.h"
#include "hml_select.h"

struct hmlGroupedPlan_t {
  int32_t count = 0;
  int64_t ntiles = 0;
  int64_t cls_off[hml::kGroupClasses] = {};   // tile list of each class
  int64_t cls_n[hml::kGroupClasses] = {};
  int cls_blocks[hml::kGroupClasses] = {};
  // per class: workgroups take blocks of 4 row tiles that share B through
  // LDS (true) or each wave takes one row tile with its own copy of B
  bool cls_shared[hml::kGroupClasses] = {};
  void *dev = nullptr;  // one allocation: problems | tiles
  hml::GroupProblem *probs = nullptr;
  hml::GroupTile *tiles = nullptr;
  size_t bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
};

namespace hml {
namespace {

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

hmlStatus record(hmlGroupedPlan_t *p, hipStream_t st) {
  for (int c = 0; c < kGroupClasses; ++c) {
    if (!p->cls_n[c]) continue;
    hipError_t e = launch_grouped(c, p->cls_shared[c], p->probs, p->tiles + p->cls_off[c],
                                  p->cls_n[c], p->cls_blocks[c], st);
    if (e != hipSuccess) return from_hip(e, "grouped kernel");
  }
  return HML_STATUS_SUCCESS;
}

}  // namespace
}  // namespace hml

using namespace hml;

extern "C" {

hmlStatus hmlSgemmGroupedCreate(hmlGroupedPlan_t **plan, int32_t count,
                                const int64_t *M, const int64_t *N,
                                const int64_t *K, const float *const *A,
                                const int64_t *lda, const float *const *B,
                                const int64_t *ldb, float *const *C,
                                const int64_t *ldc, float alpha, float beta) {
  clear_error();
  if (!plan || count <= 0 || !M || !N || !K || !A || !lda || !B || !ldb || !C || !ldc)
    return fail(HML_STATUS_INVALID_VALUE, "grouped: NULL argument or count <= 0");
  if (count > 65535) return fail(HML_STATUS_INVALID_VALUE, "grouped: at most 65535 problems");
  std::vector<GroupProblem> probs(count);
  std::vector<int64_t> tile_cost[kGroupClasses];
  std::vector<GroupTile> tiles[kGroupClasses];
  // Kernel of each class.  Shared B measured faster for N <= 32 (0.93 vs
  // 0.97 ms and 1.55 vs 1.62 ms on the 512-problem suite's classes) and
  // slower above (3.28 vs 2.96 ms: its 1 x 4 tiles hold half the rows of
  // the per-wave 2 x 4 kernel, which fits in registers only without the
  // LDS staging).  HML_GROUPED_KERNEL=wave / =shared forces one kernel
  // for every class.
  bool cls_shared[kGroupClasses] = {true, true, false};
  if (const char *kenv = getenv("HML_GROUPED_KERNEL")) {
    if (!strcmp(kenv, "wave") || !strcmp(kenv, "shared"))
      for (int c = 0; c < kGroupClasses; ++c) cls_shared[c] = kenv[0] == 's';
  }
  for (int32_t i = 0; i < count; ++i) {
    if (M[i] <= 0 || N[i] <= 0 || K[i] <= 0)
      return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d has a zero extent", i);
    if (lda[i] < K[i] || ldb[i] < N[i] || ldc[i] < N[i])
      return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d leading dimension too small", i);
    if (!A[i] || !B[i] || !C[i]) return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d NULL", i);
    if (N[i] > M[i])
      return fail(HML_STATUS_UNSUPPORTED,
                  "grouped: problem %d has N > M (only tall problems are grouped)", i);
    const int cls = N[i] <= 16 ? 0 : (N[i] <= 32 ? 1 : 2);
    int rt, ct;
    grouped_class_shape(cls, &rt, &ct);
    if (cls_shared[cls]) rt = grouped_shared_rt(cls);
    // the tile body addresses A and B with 31-bit buffer offsets
    if (((16.0 * rt - 1) * lda[i] + K[i] + 4) * 4.0 >= 2147483647.0 ||
        ((K[i] + 64.0) * ldb[i] + 16.0 * ct) * 4.0 >= 2147483647.0)
      return fail(HML_STATUS_UNSUPPORTED, "grouped: problem %d strides too large", i);
    GroupProblem &p = probs[i];
    p.A = A[i]; p.lda = lda[i]; p.B = B[i]; p.ldb = ldb[i]; p.C = C[i]; p.ldc = ldc[i];
    p.M = M[i]; p.N = N[i]; p.K = K[i];
    p.read_c = beta != 0.0f ? 1 : 0;
    p.alpha = alpha; p.beta = beta;
    const int64_t rows = 16 * rt * (cls_shared[cls] ? 4 : 1);  // rows of one work unit
    const int64_t nrb = ceil_div(M[i], rows);
    const int64_t passes = ceil_div(ceil_div(N[i], 16), ct);
    // cost of one tile: A bytes per column pass plus C bytes, in 4-byte words
    const int64_t cost = passes * rows * K[i] + rows * N[i] * (1 + p.read_c) + 64;
    for (int64_t rb = 0; rb < nrb; ++rb) {
      tiles[cls].push_back(GroupTile{i, 0, rb});
      tile_cost[cls].push_back(cost);
    }
  }
  hmlGroupedPlan_t *pl = new hmlGroupedPlan_t();
  pl->count = count;
  for (int c = 0; c < kGroupClasses; ++c) pl->cls_shared[c] = cls_shared[c];
  std::vector<GroupTile> sorted;
  for (int c = 0; c < kGroupClasses; ++c) {
    // longest-first order (stable: equal costs keep problem order)
    std::vector<int64_t> order(tiles[c].size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = (int64_t)t;
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return tile_cost[c][a] > tile_cost[c][b]; });
    pl->cls_off[c] = (int64_t)sorted.size();
    pl->cls_n[c] = (int64_t)order.size();
    for (int64_t t : order) sorted.push_back(tiles[c][t]);
    // persistent grid: the workgroups the device holds at once (4 waves
    // each), never more than needed to give every wave a tile
    int resident = grouped_class_blocks(c, cls_shared[c]);
    if (resident <= 0) resident = 512;
    pl->cls_blocks[c] = (int)std::max<int64_t>(
        1, std::min<int64_t>(resident, cls_shared[c] ? pl->cls_n[c] : ceil_div(pl->cls_n[c], 4)));
  }
  pl->ntiles = (int64_t)sorted.size();
  const size_t o_probs = 0;
  const size_t o_tiles = align256(o_probs + sizeof(GroupProblem) * count);
  pl->bytes = align256(o_tiles + sizeof(GroupTile) * std::max<size_t>(1, sorted.size()));
  hipError_t e = hipMalloc(&pl->dev, pl->bytes);
  if (e != hipSuccess) {
    const size_t b = pl->bytes;
    delete pl;
    return fail(HML_STATUS_ALLOC_FAILED, "grouped plan hipMalloc(%zu): %s", b,
                hipGetErrorString(e));
  }
  char *base = (char *)pl->dev;
  pl->probs = (GroupProblem *)(base + o_probs);
  pl->tiles = (GroupTile *)(base + o_tiles);
  e = hipMemcpy(pl->probs, probs.data(), sizeof(GroupProblem) * count, hipMemcpyHostToDevice);
  if (e == hipSuccess && !sorted.empty())
    e = hipMemcpy(pl->tiles, sorted.data(), sizeof(GroupTile) * sorted.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(pl->dev);
    delete pl;
    return from_hip(e, "grouped plan upload");
  }
  *plan = pl;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemmGroupedRun(hmlGroupedPlan_t *plan, int32_t use_graph,
                             void *stream) {
  clear_error();
  if (!plan) return fail(HML_STATUS_INVALID_VALUE, "plan is NULL");
  hipStream_t st = (hipStream_t)stream;
  if (!use_graph) return record(plan, st);
  if (!plan->exec) {
    // Capture on a private stream (the null stream cannot be captured).
    // Any failure leaves the plan as before: no stream, graph or exec.
    hipStream_t cs = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t x = nullptr;
    hipError_t e = hipStreamCreate(&cs);
    if (e != hipSuccess) return from_hip(e, "hipStreamCreate");
    hmlStatus s = HML_STATUS_SUCCESS;
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      s = from_hip(e, "hipStreamBeginCapture");
    } else {
      s = record(plan, cs);
      e = hipStreamEndCapture(cs, &g);
      if (s == HML_STATUS_SUCCESS && e != hipSuccess) s = from_hip(e, "hipStreamEndCapture");
      if (s == HML_STATUS_SUCCESS && !g)
        s = fail(HML_STATUS_HIP_ERROR, "hipStreamEndCapture returned no graph");
    }
    if (s == HML_STATUS_SUCCESS) {
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        x = nullptr;
        s = from_hip(e, "hipGraphInstantiate");
      }
    }
    if (s != HML_STATUS_SUCCESS) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipStreamDestroy(cs);
      return s;
    }
    plan->graph_stream = cs;
    plan->graph = g;
    plan->exec = x;
  }
  return from_hip(hipGraphLaunch(plan->exec, st), "hipGraphLaunch");
}

hmlStatus hmlSgemmGroupedInfo(hmlGroupedPlan_t *plan, int64_t *ntiles,
                              int32_t *nblocks, size_t *device_bytes) {
  if (!plan) return fail(HML_STATUS_INVALID_VALUE, "plan is NULL");
  if (ntiles) *ntiles = plan->ntiles;
  if (nblocks) {
    *nblocks = 0;
    for (int c = 0; c < kGroupClasses; ++c) *nblocks += plan->cls_n[c] ? plan->cls_blocks[c] : 0;
  }
  if (device_bytes) *device_bytes = plan->bytes;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemmGroupedDestroy(hmlGroupedPlan_t *plan) {
  if (!plan) return HML_STATUS_SUCCESS;
  (void)hipDeviceSynchronize();
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  if (plan->graph_stream) (void)hipStreamDestroy(plan->graph_stream);
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
  return HML_STATUS_SUCCESS;
}

}  // extern "C"
