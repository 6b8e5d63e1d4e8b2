// hml_k_noalloc.hip -- hmlSgemmExNoAlloc: hmlSgemmEx for callers that may
// not allocate device memory (code under stream capture, serving loops,
// harnesses that forbid hipMalloc).  The split-K / split-row partial sums
// that hmlSgemmEx keeps in a caller or library workspace live here in a
// fixed device array of this code object.
//
// Every call on one device shares that device's array: two calls that
// overlap in time on one device (different streams, or threads without
// ordering) race on it.  No kernel reads the array before writing it in the
// same call, so nothing in it needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "hml_runtime.h"

namespace {

constexpr size_t kScratchBytes = HML_NOALLOC_SCRATCH_BYTES;

// Each device the library runs on holds its own copy until the process
// ends.  In host code `scratch` names a host-side stand-in: kernels get the
// per-device address from hipGetSymbolAddress instead.
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

std::mutex g_scratch_mutex;
std::vector<void *> g_scratch_addr;  // device address of `scratch`, per device

}  // namespace

namespace hml {

// Resolves the current device's copy of the array once; later calls return
// the cached address.
hmlStatus noalloc_scratch(void **addr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return from_hip(e, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return from_hip(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(HML_STATUS_HIP_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_addr[dev] = p;
  }
  *addr = g_scratch_addr[dev];
  return HML_STATUS_SUCCESS;
}

}  // namespace hml

using namespace hml;

extern "C" {

hmlStatus hmlSgemmExNoAlloc(hmlOrder order, hmlOp transA, hmlOp transB,
                            int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, float beta, float *C, int64_t ldc,
                            uint32_t flags, void *stream) {
  clear_error();
  void *ws = nullptr;
  hmlStatus s = noalloc_scratch(&ws);
  if (s != HML_STATUS_SUCCESS) return s;
  s = hmlSgemmEx(order, transA, transB, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc,
                 flags, ws, kScratchBytes, stream);
  if (s == HML_STATUS_WORKSPACE_TOO_SMALL)
    return fail(s,
                "hmlSgemmExNoAlloc: %s; its fixed device scratch holds %zu bytes "
                "(hmlSgemmEx with a workspace runs this problem)",
                hmlLastError(), kScratchBytes);
  return s;
}

}  // extern "C"
