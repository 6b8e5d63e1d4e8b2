// hml_runtime.cpp -- error state, device/memory helpers of the C ABI and the
// library-owned workspace.
//
// The reference path reports every failure with fprintf + exit(1)
// (src/hml_utils.cu:75-82 hmlHandleError, src/hml_sgemm.cu:189-199) and
// allocates/frees device buffers inside every call (src/hml_utils.cu:147-230).
// The extended ABI returns status codes instead and keeps one grow-only
// scratch buffer per (device, stream).
#include "hml_runtime.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "hml_internal.h"

namespace hml {

namespace {
thread_local std::string g_last_error;

struct DeviceScratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mutex;

std::mutex g_flush_mutex;
std::vector<DeviceScratch> g_flush;
}  // namespace

hmlStatus fail(hmlStatus s, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return s;
}

hmlStatus from_hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return HML_STATUS_SUCCESS;
  return fail(HML_STATUS_HIP_ERROR, "%s: %s", what, hipGetErrorString(e));
}

void clear_error() { g_last_error.clear(); }

int verbosity() {
  static int v = [] {
    const char *s = getenv("HML_VERBOSE");
    return s ? atoi(s) : 1;
  }();
  return v;
}

[[noreturn]] void compat_die(const char *where, hmlStatus s) {
  fprintf(stderr, "%s: %s (%s)\n", where, hmlStatusString(s),
          g_last_error.c_str());
  exit(1);
}

namespace {

// One workspace slot per (device, stream).  The buffer is shared between
// the slot and the leases handed out on it; the deleter runs when the last
// of them lets go, which for a replaced buffer is the end of the last call
// that was still using it.
struct WsSlot {
  std::shared_ptr<void> buf;
  size_t bytes = 0;
  uint64_t last_use = 0;
};
struct WsKey {
  int dev;
  uintptr_t stream;
  bool operator<(const WsKey &o) const {
    return dev != o.dev ? dev < o.dev : stream < o.stream;
  }
};
std::map<WsKey, WsSlot> g_ws_slots;
uint64_t g_ws_clock = 0;
// Streams come and go; past this many slots the least recently used idle
// one is dropped (its buffer is freed once no lease holds it).
constexpr size_t kMaxWsSlots = 16;

void free_device_buffer(void *p) {
  // hipFree performs an implicit device synchronisation, so every kernel
  // queued on this buffer has finished before it is released.
  if (p) (void)hipFree(p);
}

}  // namespace

hmlStatus library_workspace(size_t bytes, hipStream_t st, WsLease *lease) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return from_hip(e, "hipGetDevice");
  std::shared_ptr<void> retired;  // released after the mutex
  std::lock_guard<std::mutex> lk(g_ws_mutex);
  const WsKey key{dev, (uintptr_t)st};
  auto it = g_ws_slots.find(key);
  if (it == g_ws_slots.end()) {
    if (g_ws_slots.size() >= kMaxWsSlots) {
      auto lru = g_ws_slots.begin();
      for (auto j = g_ws_slots.begin(); j != g_ws_slots.end(); ++j)
        if (j->second.last_use < lru->second.last_use) lru = j;
      retired = std::move(lru->second.buf);
      g_ws_slots.erase(lru);
    }
    it = g_ws_slots.emplace(key, WsSlot{}).first;
  }
  WsSlot &w = it->second;
  w.last_use = ++g_ws_clock;
  if (w.bytes < bytes || !w.buf) {
    const size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
    void *p = nullptr;
    e = hipMalloc(&p, want);
    if (e != hipSuccess)
      return fail(HML_STATUS_ALLOC_FAILED, "workspace hipMalloc(%zu): %s", want,
                  hipGetErrorString(e));
    // The old buffer lives on in the leases still holding it.
    retired = std::move(w.buf);
    w.buf = std::shared_ptr<void>(p, free_device_buffer);
    w.bytes = want;
  }
  lease->buf = w.buf;
  return HML_STATUS_SUCCESS;
}

// Device copies of the host operands of the reference-compatible entries:
// one grow-only allocation per device (A | B | C regions, 256-byte aligned),
// reused across calls instead of the reference's allocate / free per call.
namespace {
std::mutex g_compat_mutex;
std::vector<DeviceScratch> g_compat;
}  // namespace

std::mutex &compat_mutex() { return g_compat_mutex; }

hmlStatus compat_buffers(size_t a_bytes, size_t b_bytes, size_t c_bytes,
                         float **a, float **b, float **c) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return from_hip(e, "hipGetDevice");
  auto al = [](size_t x) { return (x + 255) / 256 * 256; };
  const size_t need = al(a_bytes) + al(b_bytes) + al(c_bytes);
  if ((int)g_compat.size() <= dev) g_compat.resize(dev + 1);
  DeviceScratch &w = g_compat[dev];
  if (w.bytes < need) {
    if (w.ptr) {
      e = hipFree(w.ptr);
      w.ptr = nullptr;
      w.bytes = 0;
      if (e != hipSuccess) return from_hip(e, "hipFree(staging)");
    }
    e = hipMalloc(&w.ptr, need > 0 ? need : 256);
    if (e != hipSuccess) {
      w.ptr = nullptr;
      return fail(HML_STATUS_ALLOC_FAILED, "staging hipMalloc(%zu): %s", need,
                  hipGetErrorString(e));
    }
    w.bytes = need;
  }
  char *base = (char *)w.ptr;
  *a = (float *)base;
  *b = (float *)(base + al(a_bytes));
  *c = (float *)(base + al(a_bytes) + al(b_bytes));
  return HML_STATUS_SUCCESS;
}

// Staged host <-> device copies.  The reference copies its host operands
// with one cudaMemcpy each (src/hml_utils.cu:147-230).  A pageable array
// cannot be read by the DMA engines, and registering an existing host
// allocation is not used here, so the bytes pass through pinned memory
// anyway; doing that staging here lets several host threads copy into one
// pinned chunk while the DMA engine moves the previous one.
namespace {
constexpr size_t kStageChunk = (size_t)32 << 20;  // bytes per pinned buffer
constexpr size_t kStageMin = (size_t)4 << 20;     // smaller copies: hipMemcpy

// Persistent helper threads that split one memcpy between themselves and
// the calling thread.  Created on first use and never destroyed (they wait
// on a condition variable until the process ends).
class CopyPool {
 public:
  explicit CopyPool(int helpers) {
    for (int i = 0; i < helpers; ++i) th_.emplace_back([this, i] { run(i + 1); });
  }
  void copy(void *dst, const void *src, size_t n) {
    {
      std::lock_guard<std::mutex> l(m_);
      dst_ = (char *)dst;
      src_ = (const char *)src;
      n_ = n;
      pending_ = (int)th_.size();
      ++gen_;
    }
    cv_.notify_all();
    slice(0);
    std::unique_lock<std::mutex> l(m_);
    done_.wait(l, [&] { return pending_ == 0; });
  }

 private:
  void slice(int i) const {
    const size_t parts = th_.size() + 1;
    const size_t per = ((n_ + parts - 1) / parts + 255) & ~(size_t)255;
    const size_t b = std::min(n_, per * (size_t)i), e = std::min(n_, b + per);
    if (e > b) memcpy(dst_ + b, src_ + b, e - b);
  }
  void run(int i) {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> l(m_);
        cv_.wait(l, [&] { return gen_ != seen; });
        seen = gen_;
      }
      slice(i);
      std::lock_guard<std::mutex> l(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t n_ = 0;
};

CopyPool &copy_pool() {
  static CopyPool *p = [] {
    // host threads per staged chunk (the caller included); a handful
    // saturate one PCIe link
    int t = 6;
    if (const char *s = getenv("HML_COPY_THREADS")) t = atoi(s);
    const int hw = (int)std::thread::hardware_concurrency();
    if (hw > 0) t = std::min(t, hw);
    return new CopyPool(std::max(1, std::min(t, 32)) - 1);
  }();
  return *p;
}

struct HostStage {
  void *pinned[2] = {nullptr, nullptr};
  hipStream_t st = nullptr;
  hipEvent_t done[2] = {nullptr, nullptr};
  bool unavailable = false;  // set-up failed once: plain copies from then on
};
std::vector<HostStage> g_stage;  // per device, guarded by g_compat_mutex

bool staging_off() {
  const char *s = getenv("HML_HOST_STAGING");
  return s && (!strcmp(s, "off") || !strcmp(s, "0"));
}

hmlStatus host_stage(HostStage **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return from_hip(e, "hipGetDevice");
  if ((int)g_stage.size() <= dev) g_stage.resize(dev + 1);
  HostStage &h = g_stage[dev];
  if (h.unavailable) return fail(HML_STATUS_ALLOC_FAILED, "pinned staging unavailable");
  if (!h.st) {
    hipStream_t st = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) {
      h.unavailable = true;
      return from_hip(e, "staging stream");
    }
    HostStage n;
    n.st = st;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipHostMalloc(&n.pinned[i], kStageChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&n.done[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      for (int i = 0; i < 2; ++i) {
        if (n.pinned[i]) (void)hipHostFree(n.pinned[i]);
        if (n.done[i]) (void)hipEventDestroy(n.done[i]);
      }
      (void)hipStreamDestroy(st);
      h.unavailable = true;
      return fail(HML_STATUS_ALLOC_FAILED, "pinned staging buffers: %s", hipGetErrorString(e));
    }
    h = n;
  }
  *out = &h;
  return HML_STATUS_SUCCESS;
}
}  // namespace

hmlStatus compat_h2d(void *dst, const void *src, size_t bytes) {
  HostStage *h = nullptr;
  // no pinned memory (or staging switched off): one plain copy
  if (bytes < kStageMin || staging_off() || host_stage(&h) != HML_STATUS_SUCCESS) {
    clear_error();
    return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
  hipError_t e = hipSuccess;
  size_t i = 0;
  for (size_t off = 0; off < bytes && e == hipSuccess; off += kStageChunk, ++i) {
    const int b = (int)(i & 1);
    const size_t n = std::min(kStageChunk, bytes - off);
    // the DMA read of this buffer's previous chunk must be done
    if (i >= 2 && (e = hipEventSynchronize(h->done[b])) != hipSuccess) break;
    copy_pool().copy(h->pinned[b], (const char *)src + off, n);
    e = hipMemcpyAsync((char *)dst + off, h->pinned[b], n, hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess) e = hipEventRecord(h->done[b], h->st);
  }
  const hipError_t e2 = hipStreamSynchronize(h->st);
  return from_hip(e != hipSuccess ? e : e2, "staged copy H2D");
}

hmlStatus compat_d2h(void *dst, const void *src, size_t bytes) {
  HostStage *h = nullptr;
  if (bytes < kStageMin || staging_off() || host_stage(&h) != HML_STATUS_SUCCESS) {
    clear_error();
    return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
  }
  const size_t nchunk = (bytes + kStageChunk - 1) / kStageChunk;
  auto issue = [&](size_t i) {
    const size_t off = i * kStageChunk, n = std::min(kStageChunk, bytes - off);
    hipError_t e = hipMemcpyAsync(h->pinned[i & 1], (const char *)src + off, n,
                                  hipMemcpyDeviceToHost, h->st);
    return e == hipSuccess ? hipEventRecord(h->done[i & 1], h->st) : e;
  };
  hipError_t e = issue(0);
  for (size_t i = 0; i < nchunk && e == hipSuccess; ++i) {
    // chunk i + 1 lands in the other buffer, whose chunk i - 1 the host
    // has already copied out
    if (i + 1 < nchunk && (e = issue(i + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(h->done[i & 1])) != hipSuccess) break;
    const size_t off = i * kStageChunk;
    copy_pool().copy((char *)dst + off, h->pinned[i & 1], std::min(kStageChunk, bytes - off));
  }
  const hipError_t e2 = hipStreamSynchronize(h->st);
  return from_hip(e != hipSuccess ? e : e2, "staged copy D2H");
}

}  // namespace hml

using namespace hml;

extern "C" {

const char *hmlStatusString(hmlStatus s) {
  switch (s) {
    case HML_STATUS_SUCCESS: return "success";
    case HML_STATUS_INVALID_VALUE: return "invalid value";
    case HML_STATUS_HIP_ERROR: return "HIP runtime error";
    case HML_STATUS_NO_KERNEL: return "no kernel for this variant";
    case HML_STATUS_ALLOC_FAILED: return "allocation failed";
    case HML_STATUS_IO_ERROR: return "I/O error";
    case HML_STATUS_UNSUPPORTED: return "unsupported arguments";
    case HML_STATUS_WORKSPACE_TOO_SMALL: return "workspace too small";
    case HML_STATUS_VALIDATION_FAILED: return "validation failed";
    case HML_STATUS_PARSE_ERROR: return "parse error";
  }
  return "unknown status";
}

const char *hmlLastError(void) { return g_last_error.c_str(); }

hmlStatus hmlInit(void) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return from_hip(e, "hipGetDeviceCount");
  if (n <= 0) return fail(HML_STATUS_HIP_ERROR, "no GPU device");
  // set-up of hmlSgemmExNoAlloc's fixed scratch on the current device
  void *scratch = nullptr;
  return noalloc_scratch(&scratch);
}

hmlStatus hmlSetDevice(int device) {
  return from_hip(hipSetDevice(device), "hipSetDevice");
}

int hmlDeviceCount(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

hmlStatus hmlDeviceInfo(int device, char *name, int32_t *cus,
                        int64_t *l2_bytes, int64_t *hbm_bytes,
                        int32_t *clock_khz) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return from_hip(e, "hipGetDeviceProperties");
  if (name) {
    snprintf(name, 64, "%s", prop.gcnArchName);
  }
  if (cus) *cus = prop.multiProcessorCount;
  if (l2_bytes) *l2_bytes = prop.l2CacheSize;
  if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
  if (clock_khz) *clock_khz = prop.clockRate;
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlDeviceMalloc(void **ptr, size_t bytes) {
  if (!ptr) return fail(HML_STATUS_INVALID_VALUE, "ptr is NULL");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess)
    return fail(HML_STATUS_ALLOC_FAILED, "hipMalloc(%zu): %s", bytes,
                hipGetErrorString(e));
  return HML_STATUS_SUCCESS;
}

hmlStatus hmlDeviceFree(void *ptr) { return from_hip(hipFree(ptr), "hipFree"); }

hmlStatus hmlMemcpyHtoD(void *dst, const void *src, size_t bytes) {
  return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
hmlStatus hmlMemcpyDtoH(void *dst, const void *src, size_t bytes) {
  return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
hmlStatus hmlMemcpyDtoD(void *dst, const void *src, size_t bytes) {
  return from_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}
hmlStatus hmlMemsetD8(void *dst, int value, size_t bytes) {
  return from_hip(hipMemset(dst, value, bytes), "hipMemset");
}
hmlStatus hmlDeviceSynchronize(void) {
  return from_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}
hmlStatus hmlStreamCreate(void **stream) {
  hipStream_t s;
  hipError_t e = hipStreamCreate(&s);
  if (e != hipSuccess) return from_hip(e, "hipStreamCreate");
  *stream = s;
  return HML_STATUS_SUCCESS;
}
hmlStatus hmlStreamDestroy(void *stream) {
  return from_hip(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}
hmlStatus hmlStreamSynchronize(void *stream) {
  return from_hip(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}
hmlStatus hmlEventCreate(void **event) {
  hipEvent_t ev;
  hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) return from_hip(e, "hipEventCreate");
  *event = ev;
  return HML_STATUS_SUCCESS;
}
hmlStatus hmlEventDestroy(void *event) {
  return from_hip(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}
hmlStatus hmlEventRecord(void *event, void *stream) {
  return from_hip(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}
hmlStatus hmlEventSynchronize(void *event) {
  return from_hip(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}
hmlStatus hmlEventElapsedMs(float *ms, void *start, void *stop) {
  return from_hip(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
                  "hipEventElapsedTime");
}

hmlStatus hmlReadProbe(const void *src, size_t bytes, void *sink, void *stream) {
  if (!src || !sink || bytes % 16 != 0)
    return fail(HML_STATUS_INVALID_VALUE, "read probe: src, sink and a multiple of 16 bytes");
  return from_hip(launch_read_probe(src, bytes, (uint32_t *)sink, (hipStream_t)stream),
                  "read probe");
}

hmlStatus hmlFlushCaches(void *stream) {
  // 512 MiB of writes evicts the 256 MB MALL and every XCD's L2.
  const size_t bytes = (size_t)512 << 20;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return from_hip(e, "hipGetDevice");
  void *p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_flush_mutex);
    if ((int)g_flush.size() <= dev) g_flush.resize(dev + 1);
    if (!g_flush[dev].ptr) {
      e = hipMalloc(&g_flush[dev].ptr, bytes);
      if (e != hipSuccess) {
        g_flush[dev].ptr = nullptr;
        return fail(HML_STATUS_ALLOC_FAILED, "flush buffer: %s", hipGetErrorString(e));
      }
      g_flush[dev].bytes = bytes;
    }
    p = g_flush[dev].ptr;
  }
  return from_hip(launch_fill((float *)p, bytes / sizeof(float), 0.f,
                              (hipStream_t)stream),
                  "flush fill");
}

}  // extern "C"
