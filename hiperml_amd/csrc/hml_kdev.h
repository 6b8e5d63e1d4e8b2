// hml_kdev.h -- device-side helpers shared by the gfx950 kernels (included
// only by .hip sources).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hml {

typedef float f4 __attribute__((ext_vector_type(4)));

// Raw buffer descriptor (gfx9 format word) of `bytes` bytes at `base`.  A
// load whose lane offset is >= the size returns zero and reads nothing; the
// wave-uniform soffset is not part of that check, so every lane-dependent
// part of an address goes into the lane offset.
constexpr int kBufNT = 2;                 // cache policy bit: non-temporal
constexpr int32_t kRsrcMax = 0x7fffffff;  // size of the operand descriptors
constexpr uint32_t kOOB = 0x80000000u;    // lane offset past every descriptor
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base,
                                                           int32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, bytes, 0x00020000);
}
// 16-byte load at any 4-byte-aligned offset: buffer_load_dwordx4 needs only
// dword alignment on MI355X (checked bit-exact with odd leading dimensions
// by the parity tests).
template <int AUX>
__device__ __forceinline__ f4 buf_load4(__amdgpu_buffer_rsrc_t rs, uint32_t voff,
                                        int32_t soff) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff, AUX));
}
// 16-byte buffer load straight into LDS: lane l's 16 bytes land at
// lds + 16 l (buffer_load_dwordx4 ... lds; the LDS base goes through M0).
// The builtin exists only for the device target, so the host pass of a
// kernel that instantiates this sees an empty body.
template <int AUX>
__device__ __forceinline__ void buf_load_lds16(__amdgpu_buffer_rsrc_t rs, float *lds,
                                               uint32_t voff, int32_t soff) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void *)lds, 16,
                                           voff, soff, 0, AUX);
#endif
}
typedef float f2 __attribute__((ext_vector_type(2)));
template <int AUX = 0>
__device__ __forceinline__ f2 buf_load2(__amdgpu_buffer_rsrc_t rs, uint32_t voff, int32_t soff) {
  return __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, AUX));
}
template <int AUX = 0>
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t rs, uint32_t voff,
                                           int32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, AUX));
}

}  // namespace hml
