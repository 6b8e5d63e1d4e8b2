// hml_dense.h -- dense (large-N) SGEMM kernels: plain FP32 MFMA and
// split-precision bf16 MFMA (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hml {

constexpr int kDenseTileM = 128;  // FP32 kernel block tile
constexpr int kDenseTileN = 128;
constexpr int kDenseTileK = 32;   // padding granularity of K
constexpr int kSplitTileM = 256;  // 16-bit planes' row padding (largest block tile)
constexpr int kSplitTileN = 128;
constexpr int kSplitKBlock = 32;  // k-block of the bf16 plane layout
// padded K up to which the 16-bit kernels use 128-row block tiles
constexpr int64_t kDenseShortK = 1024;

struct DenseParams {
  // S == 0: FP32 [Mp][Kp] / [Np][Kp] (B transposed, K contiguous);
  // S > 0:  bf16 planes, k-blocked [S][Kp/32][Mp][32] / [S][Kp/32][Np][32]
  const void *Ap;
  const void *Bp;
  float *C;
  int64_t sCm, sCn;
  int64_t M, N, Mp, Np, Kp;
  float alpha, beta;
  int32_t read_c;
  // scaled fp16 split only: per-row maxima (FP32 bits) of A and of B^T,
  // from which both the split and the epilogue derive the scale exponents
  const uint32_t *rmaxA;
  const uint32_t *rmaxB;
};

// S: 0 = FP32 copy, 2 = hi/lo bf16, 3 = hi/mid/lo bf16.
hipError_t launch_split(const float *src, int64_t sR, int64_t sK, int64_t rows,
                        int64_t K, int64_t rowsP, int64_t Kp, int S, void *dst,
                        hipStream_t st);
// S: 0 FP32, 2 bf16x3, 3 bf16x6, 4 scaled f16x3.
hipError_t launch_dense(int S, const DenseParams &p, hipStream_t st);
// rowmax[r] = max_k |X(r,k)| as FP32 bits; rowmax must be zero on entry.
hipError_t launch_rowmax(const float *src, int64_t sR, int64_t sK, int64_t rows,
                         int64_t K, uint32_t *rowmax, hipStream_t st);
// Scaled fp16 hi / lo planes (k-blocked, 2 planes) from the row maxima.
hipError_t launch_split_f16(const float *src, int64_t sR, int64_t sK, int64_t rows,
                            int64_t K, int64_t rowsP, int64_t Kp, const uint32_t *rowmax,
                            void *dst, hipStream_t st);

}  // namespace hml
