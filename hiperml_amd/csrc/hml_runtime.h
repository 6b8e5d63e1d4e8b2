// hml_runtime.h -- host-side runtime shared by the C-ABI translation units:
// error reporting, per-device workspace, argument helpers.
#pragma once

#include <memory>
#include <mutex>

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "../../include/hml_ext.h"

namespace hml {

// Records `msg` as the thread's last error and returns `s`.
hmlStatus fail(hmlStatus s, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));
hmlStatus from_hip(hipError_t e, const char *what);
void clear_error();

// A hold on a library-owned workspace buffer.  The library keeps one
// grow-only buffer per (device, stream): calls on one stream are ordered by
// the stream, so they may share it; calls on different streams never do.
// A buffer replaced by a larger one is freed only when the last lease on it
// ends (hipFree then waits for the device, so kernels queued under the lease
// have finished).  Keep the lease alive until the kernels that use ptr() are
// queued.
struct WsLease {
  std::shared_ptr<void> buf;
  void *ptr() const { return buf.get(); }
};
hmlStatus library_workspace(size_t bytes, hipStream_t st, WsLease *lease);

// Device address of the current device's fixed scratch array of
// hmlSgemmExNoAlloc (resolved once per device).
hmlStatus noalloc_scratch(void **addr);

// Splits a workspace into 256-byte aligned pieces.
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Verbosity of the reference-style "Info:" selection lines (HML_VERBOSE).
int verbosity();

// Compat-mode failure: print and exit(1), as the reference does
// (src/hml_sgemm.cu:189-199, src/hml_utils.cu:75-82).
[[noreturn]] void compat_die(const char *where, hmlStatus s);

// Device copies of the compat entries' host operands: A | B | C regions of
// one grow-only per-device allocation.  Hold compat_mutex() while using them
// (one compat call at a time per process).
std::mutex &compat_mutex();
hmlStatus compat_buffers(size_t a_bytes, size_t b_bytes, size_t c_bytes,
                         float **a, float **b, float **c);
// Synchronous copies between a pageable host array and device memory for
// the compat entries (hold compat_mutex()): large ones move in chunks
// through two pinned staging buffers, host threads filling (draining) one
// while the DMA engine drains (fills) the other.  HML_HOST_STAGING=off, or
// pinned memory that cannot be had, makes them plain hipMemcpy calls.
hmlStatus compat_h2d(void *dst, const void *src, size_t bytes);
hmlStatus compat_d2h(void *dst, const void *src, size_t bytes);

}  // namespace hml
