// hml_runtime.h -- host-side runtime shared by the C-ABI translation units:
// error reporting, per-device workspace, argument helpers.
#pragma once

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

// Library-owned per-device scratch, grown on demand (hipFree synchronises
// the device, so growth never frees memory a queued kernel still uses).
hmlStatus library_workspace(size_t bytes, void **ptr);

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

}  // namespace hml
