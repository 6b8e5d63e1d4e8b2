// hml_grouped.cpp -- grouped (batched) skinny SGEMM: a plan holds the device
// descriptors of many independent row-major problems and, per tile class
// (by N), a kernel kind and a balanced static work list laid out by XCD; running it
// launches one persistent kernel per non-empty class, optionally replayed
// from a captured hipGraph.  hmlSgemmGroupedAutotune picks each class's
// kind by timing all of them on the plan's own problems.
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
  std::vector<hml::GroupProblem> hprobs;  // host copy of the descriptors
  std::vector<int> pcls;                  // tile class of each problem
  int64_t ntiles = 0;                     // work-list entries, all classes
  int64_t cls_off[hml::kGroupClasses] = {};  // region of each class in `tiles`
  int64_t cls_cap[hml::kGroupClasses] = {};  // entries the region holds
  int64_t cls_n[hml::kGroupClasses] = {};    // list length of the current kind
  int64_t cls_real[hml::kGroupClasses] = {};  // its entries (the rest are holes)
  int cls_blocks[hml::kGroupClasses] = {};
  int cls_kind[hml::kGroupClasses] = {};
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

static_assert(kGroupClasses == HML_GROUPED_CLASSES, "tile classes of the ABI");
static_assert(kGroupWave2 == HML_GROUPED_WAVE2 && kGroupWave1 == HML_GROUPED_WAVE1 &&
                  kGroupShared == HML_GROUPED_SHARED && kGroupKinds == 3,
              "kernel kinds of the ABI");

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

hmlStatus record(hmlGroupedPlan_t *p, hipStream_t st) {
  for (int c = 0; c < kGroupClasses; ++c) {
    if (!p->cls_n[c]) continue;
    hipError_t e = launch_grouped(c, p->cls_kind[c], p->probs, p->tiles + p->cls_off[c],
                                  p->cls_n[c], p->cls_blocks[c], st);
    if (e != hipSuccess) return from_hip(e, "grouped kernel");
  }
  return HML_STATUS_SUCCESS;
}

void drop_graph(hmlGroupedPlan_t *p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->graph_stream) (void)hipStreamDestroy(p->graph_stream);
  p->exec = nullptr;
  p->graph = nullptr;
  p->graph_stream = nullptr;
}

// Entries of class c for kernel kind `kind`: every (problem, row block),
// longest first (stable: problems of equal cost keep their order), with
// its cost: the A words of its row block times the column passes, plus its
// C words.
void class_entries(const hmlGroupedPlan_t *p, int c, int kind, std::vector<GroupTile> *tiles,
                   std::vector<int64_t> *cost) {
  int rt, ct;
  grouped_kind_shape(c, kind, &rt, &ct);
  const int64_t rows = grouped_unit_rows(c, kind);
  std::vector<std::pair<int64_t, int32_t>> order;  // (cost, problem)
  for (int32_t i = 0; i < p->count; ++i) {
    if (p->pcls[i] != c) continue;
    const GroupProblem &g = p->hprobs[i];
    const int64_t passes = ceil_div(ceil_div(g.N, 16), ct);
    order.emplace_back(passes * rows * g.K + rows * g.N * (1 + g.read_c) + 64, i);
  }
  std::stable_sort(order.begin(), order.end(),
                   [](const std::pair<int64_t, int32_t> &a, const std::pair<int64_t, int32_t> &b) {
                     return a.first > b.first;
                   });
  tiles->clear();
  cost->clear();
  for (const auto &o : order) {
    const int64_t nrb = ceil_div(p->hprobs[o.second].M, rows);
    for (int64_t rb = 0; rb < nrb; ++rb) {
      tiles->push_back(GroupTile{o.second, 0, rb});
      cost->push_back(o.first);
    }
  }
}

// Lays the longest-first entries out for the kernels' static schedule, in
// which scheduling unit u (a wave, or a workgroup in the shared-B kind) of
// U runs list positions u, u + U, u + 2U, ...  Plain longest-first order
// deals round r's U entries to units 0..U-1 every round, so the same units
// keep receiving the larger entry of each round: on the 512-problem suite
// the busiest unit carried 1.08-1.18x the mean cost.  Here each round's
// entries go, largest first, to the units with the least cost so far
// (1.00-1.01x the mean), and the units that take part in the last, partial
// round are numbered first so that those entries land at the end of the
// list.
void balance_rounds(const std::vector<GroupTile> &in, const std::vector<int64_t> &cost, int64_t U,
                    std::vector<GroupTile> *out) {
  const int64_t n = (int64_t)in.size();
  out->assign(in.begin(), in.end());
  if (U <= 1 || n <= U) return;
  const int64_t R = ceil_div(n, U);
  std::vector<int64_t> load(U, 0), unit(n), by_load(U);
  for (int64_t u = 0; u < U; ++u) by_load[u] = u;
  for (int64_t r = 0; r < R; ++r) {
    const int64_t b = r * U, m = std::min(U, n - b);
    std::stable_sort(by_load.begin(), by_load.end(),
                     [&](int64_t a, int64_t c) { return load[a] < load[c]; });
    for (int64_t i = 0; i < m; ++i) {
      unit[b + i] = by_load[i];
      load[by_load[i]] += cost[b + i];
    }
  }
  // physical index of each unit: the last round's units first
  std::vector<int64_t> phys(U, -1);
  const int64_t last = (R - 1) * U;
  int64_t next = 0;
  for (int64_t i = last; i < n; ++i) phys[unit[i]] = next++;
  for (int64_t u = 0; u < U; ++u)
    if (phys[u] < 0) phys[u] = next++;
  for (int64_t i = 0; i < n; ++i) (*out)[(i / U) * U + phys[unit[i]]] = in[i];
}

// Workgroups are dealt to the XCDs round robin: workgroup i runs on XCD
// i % kXcds, each XCD with its own L2.
constexpr int kXcds = 8;

// balance_rounds with the units grouped by XCD, so that the entries one
// XCD runs at a time come from few problems, whose B then stays in that
// XCD's L2 (dealt over all units, a round's entries put every active
// problem's B in every XCD's L2).  Round r's entries (longest first, each
// problem's row blocks together) are cut into one contiguous slice per
// XCD, sized by its units; the slices go, largest first, to the XCDs with
// the least cost so far, and each slice's entries, largest first, to the
// XCD's units with the least cost so far.  Unit u (a wave, u / per_block
// its workgroup) runs positions u, u + U, ...; in the last, partial round
// the units without an entry get holes (prob = -1, skipped by the
// kernels).  Returns false (nothing laid out) with fewer workgroups than
// XCDs or a list of one round, where the plain layout does as well.
bool balance_rounds_xcd(const std::vector<GroupTile> &in, const std::vector<int64_t> &cost,
                        int64_t U, int per_block, std::vector<GroupTile> *out) {
  const int64_t n = (int64_t)in.size();
  if (U / per_block < kXcds || n <= U) return false;
  std::vector<std::vector<int64_t>> units(kXcds);
  for (int64_t u = 0; u < U; ++u) units[(u / per_block) % kXcds].push_back(u);
  std::vector<int64_t> load(U, 0), xload(kXcds, 0);
  std::vector<int> xorder(kXcds);
  const int64_t R = ceil_div(n, U);
  out->assign((size_t)(R * U), GroupTile{-1, 0, 0});
  int64_t last = 0;  // one past the last position written
  for (int64_t r = 0; r < R; ++r) {
    for (int x = 0; x < kXcds; ++x) xorder[x] = x;
    std::stable_sort(xorder.begin(), xorder.end(),
                     [&](int a, int b) { return xload[a] < xload[b]; });
    int64_t i = r * U;
    const int64_t e = std::min(n, i + U);
    for (int k = 0; k < kXcds && i < e; ++k) {
      std::vector<int64_t> &us = units[xorder[k]];
      std::stable_sort(us.begin(), us.end(), [&](int64_t a, int64_t b) { return load[a] < load[b]; });
      const int64_t m = std::min<int64_t>((int64_t)us.size(), e - i);
      for (int64_t j = 0; j < m; ++j, ++i) {
        const int64_t pos = r * U + us[j];
        (*out)[pos] = in[i];
        load[us[j]] += cost[i];
        xload[xorder[k]] += cost[i];
        last = std::max(last, pos + 1);
      }
    }
  }
  out->resize((size_t)last);
  return true;
}

// Switches class c to kernel kind `kind`: uploads its work list and sizes
// its persistent grid.  Nothing may be running the plan.
hmlStatus set_kind(hmlGroupedPlan_t *p, int c, int kind) {
  std::vector<GroupTile> sorted, t;
  std::vector<int64_t> cost;
  class_entries(p, c, kind, &sorted, &cost);
  if ((int64_t)sorted.size() > p->cls_cap[c])
    return fail(HML_STATUS_INVALID_VALUE, "grouped: work list overflow");
  // the workgroups the device holds at once, never more than the list
  // needs (one entry per workgroup in the shared-B kind, one per wave
  // otherwise)
  const int64_t n = (int64_t)sorted.size();
  int resident = grouped_class_blocks(c, kind);
  if (resident <= 0) resident = kind == kGroupShared ? 512 : 2048 / kGroupWaveWpb;
  const int64_t need = kind == kGroupShared ? n : ceil_div(n, kGroupWaveWpb);
  const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(resident, need));
  const int per_block = kind == kGroupShared ? 1 : kGroupWaveWpb;
  const int64_t U = (int64_t)blocks * per_block;
  // the XCD layout when its holes fit the class's room, else the plain one
  // (which has none)
  if (!balance_rounds_xcd(sorted, cost, U, per_block, &t) || (int64_t)t.size() > p->cls_cap[c])
    balance_rounds(sorted, cost, U, &t);
  if (!t.empty()) {
    hipError_t e = hipMemcpy(p->tiles + p->cls_off[c], t.data(), sizeof(GroupTile) * t.size(),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) return from_hip(e, "grouped work list upload");
  }
  p->ntiles += n - p->cls_real[c];
  p->cls_real[c] = n;
  p->cls_n[c] = (int64_t)t.size();
  p->cls_kind[c] = kind;
  p->cls_blocks[c] = blocks;
  drop_graph(p);  // a captured graph holds the old launches
  return HML_STATUS_SUCCESS;
}

// Default kernel kind of each class.  Shared B measured faster for N <= 32
// on the 512-problem suite (0.84 vs 0.91 ms and 1.42 vs 1.48 ms for those
// two classes) and slower above (3.11 vs 2.90 ms for N = 33..64 in one
// class: its 1 x 4 tiles hold half the rows of the per-wave 2 x 4 kernel).
// HML_GROUPED_KERNEL = wave, wave1 or shared forces one kind for every
// class; hmlSgemmGroupedAutotune measures.
void default_kinds(int kinds[kGroupClasses]) {
  kinds[0] = kGroupShared;
  kinds[1] = kGroupShared;
  kinds[2] = kGroupWave2;
  kinds[3] = kGroupWave2;
  if (const char *k = getenv("HML_GROUPED_KERNEL")) {
    int f = -1;
    if (!strcmp(k, "wave")) f = kGroupWave2;
    if (!strcmp(k, "wave1")) f = kGroupWave1;
    if (!strcmp(k, "shared")) f = kGroupShared;
    if (f >= 0)
      for (int c = 0; c < kGroupClasses; ++c) kinds[c] = f;
  }
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
  std::vector<int> pcls(count);
  int64_t cap[kGroupClasses] = {};
  for (int32_t i = 0; i < count; ++i) {
    if (M[i] <= 0 || N[i] <= 0 || K[i] <= 0)
      return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d has a zero extent", i);
    if (lda[i] < K[i] || ldb[i] < N[i] || ldc[i] < N[i])
      return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d leading dimension too small", i);
    if (!A[i] || !B[i] || !C[i]) return fail(HML_STATUS_INVALID_VALUE, "grouped: problem %d NULL", i);
    if (N[i] > M[i])
      return fail(HML_STATUS_UNSUPPORTED,
                  "grouped: problem %d has N > M (only tall problems are grouped)", i);
    // the tile bodies address A and B with 31-bit buffer offsets from a
    // wave's first row (at most 2 row tiles of 16) and column pass (at most
    // 4 column tiles of 16)
    if ((31.0 * lda[i] + K[i] + 4) * 4.0 >= 2147483647.0 ||
        ((K[i] + 64.0) * ldb[i] + 64.0) * 4.0 >= 2147483647.0)
      return fail(HML_STATUS_UNSUPPORTED, "grouped: problem %d strides too large", i);
    GroupProblem &p = probs[i];
    p.A = A[i]; p.lda = lda[i]; p.B = B[i]; p.ldb = ldb[i]; p.C = C[i]; p.ldc = ldc[i];
    p.M = M[i]; p.N = N[i]; p.K = K[i];
    p.read_c = beta != 0.0f ? 1 : 0;
    p.alpha = alpha; p.beta = beta;
    pcls[i] = grouped_class_of(N[i]);
    cap[pcls[i]] += ceil_div(M[i], 16);  // the finest work list (16-row entries)
  }
  // room for the holes of the XCD layout's last round (under one round of
  // units: at most 16 resident waves per CU)
  for (int c = 0; c < kGroupClasses; ++c)
    if (cap[c]) cap[c] += 16384;
  hmlGroupedPlan_t *pl = new hmlGroupedPlan_t();
  pl->count = count;
  pl->hprobs = probs;
  pl->pcls = pcls;
  int64_t total = 0;
  for (int c = 0; c < kGroupClasses; ++c) {
    pl->cls_off[c] = total;
    pl->cls_cap[c] = cap[c];
    total += cap[c];
  }
  const size_t o_probs = 0;
  const size_t o_tiles = align256(o_probs + sizeof(GroupProblem) * count);
  pl->bytes = align256(o_tiles + sizeof(GroupTile) * std::max<int64_t>(1, total));
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
  hmlStatus s = from_hip(e, "grouped plan upload");
  int kinds[kGroupClasses];
  default_kinds(kinds);
  for (int c = 0; c < kGroupClasses && s == HML_STATUS_SUCCESS; ++c) s = set_kind(pl, c, kinds[c]);
  if (s != HML_STATUS_SUCCESS) {
    (void)hipFree(pl->dev);
    delete pl;
    return s;
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

hmlStatus hmlSgemmGroupedKinds(hmlGroupedPlan_t *plan, int32_t *kinds) {
  if (!plan || !kinds) return fail(HML_STATUS_INVALID_VALUE, "plan or kinds is NULL");
  for (int c = 0; c < kGroupClasses; ++c) kinds[c] = plan->cls_n[c] ? plan->cls_kind[c] : -1;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlSgemmGroupedSetKind(hmlGroupedPlan_t *plan, int32_t cls, int32_t kind) {
  clear_error();
  if (!plan) return fail(HML_STATUS_INVALID_VALUE, "plan is NULL");
  if (cls < 0 || cls >= kGroupClasses || kind < 0 || kind >= kGroupKinds)
    return fail(HML_STATUS_INVALID_VALUE, "grouped: class %d / kind %d out of range", cls, kind);
  hipError_t e = hipDeviceSynchronize();  // the old work list may still be in use
  if (e != hipSuccess) return from_hip(e, "hipDeviceSynchronize");
  return set_kind(plan, cls, kind);
}

hmlStatus hmlSgemmGroupedAutotune(hmlGroupedPlan_t *plan, int32_t reps, void *stream,
                                  float *ms) {
  clear_error();
  if (!plan) return fail(HML_STATUS_INVALID_VALUE, "plan is NULL");
  if (reps < 1 || reps > 1000) return fail(HML_STATUS_INVALID_VALUE, "reps must be 1..1000");
  hipStream_t st = (hipStream_t)stream;
  if (ms)
    for (int i = 0; i < kGroupClasses * kGroupKinds; ++i) ms[i] = -1.0f;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return from_hip(e, "hipDeviceSynchronize");
  // the trial runs overwrite C: keep a packed copy and put it back
  size_t cbytes = 0;
  for (const GroupProblem &g : plan->hprobs) cbytes += (size_t)g.M * g.N * sizeof(float);
  void *saved = nullptr;
  if ((e = hipMalloc(&saved, cbytes)) != hipSuccess)
    return fail(HML_STATUS_ALLOC_FAILED, "grouped autotune: copy of C (%zu bytes): %s", cbytes,
                hipGetErrorString(e));
  auto copy_c = [&](bool save) {
    size_t off = 0;
    hipError_t r = hipSuccess;
    for (const GroupProblem &g : plan->hprobs) {
      char *packed = (char *)saved + off;
      const size_t row = (size_t)g.N * sizeof(float), pitch = (size_t)g.ldc * sizeof(float);
      r = save ? hipMemcpy2DAsync(packed, row, g.C, pitch, row, g.M, hipMemcpyDeviceToDevice, st)
               : hipMemcpy2DAsync(g.C, pitch, packed, row, row, g.M, hipMemcpyDeviceToDevice, st);
      if (r != hipSuccess) break;
      off += row * g.M;
    }
    return r != hipSuccess ? r : hipStreamSynchronize(st);
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hmlStatus s = from_hip(copy_c(true), "grouped autotune: save C");
  const bool have_copy = s == HML_STATUS_SUCCESS;
  if (s == HML_STATUS_SUCCESS) {
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    s = from_hip(e, "hipEventCreate");
  }
  for (int c = 0; c < kGroupClasses && s == HML_STATUS_SUCCESS; ++c) {
    if (!plan->cls_n[c]) continue;
    int best = plan->cls_kind[c];
    float best_ms = 0.0f;
    for (int kind = 0; kind < kGroupKinds && s == HML_STATUS_SUCCESS; ++kind) {
      if ((s = set_kind(plan, c, kind)) != HML_STATUS_SUCCESS) break;
      auto run = [&] {
        return launch_grouped(c, kind, plan->probs, plan->tiles + plan->cls_off[c],
                              plan->cls_n[c], plan->cls_blocks[c], st);
      };
      e = run();  // warm-up
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int r = 0; r < reps && e == hipSuccess; ++r) e = run();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.0f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if ((s = from_hip(e, "grouped autotune run")) != HML_STATUS_SUCCESS) break;
      t /= (float)reps;
      if (ms) ms[c * kGroupKinds + kind] = t;
      if (kind == 0 || t < best_ms) {
        best = kind;
        best_ms = t;
      }
    }
    if (s == HML_STATUS_SUCCESS) s = set_kind(plan, c, best);
  }
  if (have_copy) {
    const hipError_t er = copy_c(false);
    if (s == HML_STATUS_SUCCESS) s = from_hip(er, "grouped autotune: restore C");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(saved);
  return s;
}

hmlStatus hmlSgemmGroupedDestroy(hmlGroupedPlan_t *plan) {
  if (!plan) return HML_STATUS_SUCCESS;
  (void)hipDeviceSynchronize();
  drop_graph(plan);
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
  return HML_STATUS_SUCCESS;
}

}  // extern "C"
