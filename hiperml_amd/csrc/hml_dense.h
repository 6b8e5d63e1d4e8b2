// hml_dense.h -- dense (large-N) SGEMM kernels: plain FP32 MFMA and
// split-precision bf16 MFMA (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hml {

constexpr int kDenseTileM = 128;
constexpr int kDenseTileN = 128;
constexpr int kDenseTileK = 32;  // padding granularity of the split planes

struct DenseParams {
  const void *Ap;  // [S][Mp][Kp] bf16 planes (S > 0) or [Mp][Kp] FP32 (S == 0)
  const void *Bp;  // [S][Np][Kp] likewise, B transposed (K contiguous)
  float *C;
  int64_t sCm, sCn;
  int64_t M, N, Mp, Np, Kp;
  float alpha, beta;
  int32_t read_c;
};

// S: 0 = FP32 copy, 2 = hi/lo bf16, 3 = hi/mid/lo bf16.
hipError_t launch_split(const float *src, int64_t sR, int64_t sK, int64_t rows,
                        int64_t K, int64_t rowsP, int64_t Kp, int S, void *dst,
                        hipStream_t st);
hipError_t launch_dense(int S, const DenseParams &p, hipStream_t st);

}  // namespace hml
