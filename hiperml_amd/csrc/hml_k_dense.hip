// hml_k_dense.hip -- dense SGEMM on gfx950 matrix cores: plain FP32 and
// split-precision bf16 (no counterpart in the reference, which has only the
// skinny FP32 kernels).
//
// Split precision (Henry, Tang & Heinecke, "Leveraging the bfloat16
// Artificial Intelligence Datatype For Higher-Precision Computations",
// ARITH 2019): every FP32 value x is split into S bf16 terms
//   h = bf16(x), m = bf16(x - h), l = bf16(x - h - m)
// (each residual is exact in FP32), and C = sum of the leading products:
//   S = 2 ("bf16x3"): h*h + h*l' + l'*h           (l' = bf16(x - h))
//   S = 3 ("bf16x6"): h*h + h*m + m*h + h*l + l*h + m*m
// on v_mfma_f32_16x16x32_bf16 with FP32 accumulation.  bf16 x bf16 products
// are exact in FP32, so the error comes from the dropped terms
// (2^-16 relative for S = 2, ~2^-24 for S = 3) plus FP32 accumulation.
//
// Data path: a split pass writes zero-padded bf16 planes in a k-blocked
// layout, Ap[S][Kp/32][Mp][32] and Bp[S][Kp/32][Np][32] (B transposed): the
// 32-wide k-step of a 256-row panel is then 16 KB of contiguous memory, read
// with whole-line 16-byte loads and no bounds checks.  Block tile 256 x 128 x
// 32, 8 waves in a 4 x 2 grid of 64 x 64 wave tiles (2 waves per SIMD), two
// LDS buffers (one barrier per k-step), global loads of step t+1 in flight
// while step t computes, LDS rows unpadded with the 16-byte chunk index
// XOR-swizzled by the row (conflict-free row-contiguous writes and operand
// reads), and an XCD-aware block order (blocks that share A/B panels run on
// the same L2).  The plain FP32 kernel keeps K-contiguous FP32 copies and
// 128 x 128 tiles.
#include "hml_internal.h"
#include "hml_dense.h"

namespace hml {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int BM = kDenseTileM, BN = kDenseTileN;   // FP32 kernel tile
constexpr int BMH = kSplitTileM, BNH = kSplitTileN;  // bf16 kernels' tile
constexpr int BKF = 16;        // k per step, FP32 kernel
constexpr int PADF = 4;        // float elements of LDS row padding

// Round-to-nearest-even FP32 -> bf16 bits (finite inputs; NaN kept quiet).
__device__ __forceinline__ uint32_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ __forceinline__ float bf16_to_float(uint32_t b) {
  return __uint_as_float(b << 16);
}

// ---- split / pack pass -------------------------------------------------------
// dst plane s (bf16, or FP32 when S == 0) of a rows x K operand X(r,k) =
// src[r*sR + k*sK], padded to rowsP x Kp.  One thread per 8 consecutive k of
// one row.  FP32 copies: the thread index runs fastest along whichever of
// r / k is contiguous in src, so the reads coalesce.  bf16 planes
// (k-blocked): thread idx writes the 8 elements at idx * 8 of each plane
// (chunk q = idx % 4 of row (idx / 4) % rowsP in k-block (idx / 4) / rowsP),
// so the writes are contiguous and the reads cover 128 bytes of a row (K
// contiguous) or 4 interleaved runs along the rows (rows contiguous).
template <int S>
__global__ void k_split(const float *__restrict__ src, int64_t sR, int64_t sK,
                        int64_t rows, int64_t K, int64_t rowsP, int64_t Kp,
                        void *__restrict__ dst, int kfast) {
  const int64_t kc_n = Kp / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rowsP * kc_n) return;
  int64_t r, kc;
  if (S > 0) {
    const int64_t rq = idx >> 2;
    const int64_t kt = rq / rowsP;
    r = rq - kt * rowsP;
    kc = kt * 4 + (idx & 3);
  } else if (kfast) {
    r = idx / kc_n;
    kc = idx - r * kc_n;
  } else {
    kc = idx / rowsP;
    r = idx - kc * rowsP;
  }
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int64_t k = kc * 8 + j;
    v[j] = (r < rows && k < K) ? src[r * sR + k * sK] : 0.f;
  }
  if (S == 0) {
    float *d = (float *)dst + r * Kp + kc * 8;
    *reinterpret_cast<f4 *>(d) = f4{v[0], v[1], v[2], v[3]};
    *reinterpret_cast<f4 *>(d + 4) = f4{v[4], v[5], v[6], v[7]};
    return;
  }
  const int64_t plane = rowsP * Kp;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    uint32_t b[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      b[j] = bf16_bits(v[j]);
      v[j] = v[j] - bf16_to_float(b[j]);  // exact residual
    }
    u4 w = {b[0] | (b[1] << 16), b[2] | (b[3] << 16), b[4] | (b[5] << 16),
            b[6] | (b[7] << 16)};
    *reinterpret_cast<u4 *>((uint16_t *)dst + s * plane + idx * 8) = w;
  }
}

// Block index -> (bm, bn): consecutive launch indices rotate over the 8
// XCDs, so first regroup them per XCD, then walk each XCD's share in
// column groups of 8 block-rows so its blocks reuse A and B panels in L2.
__device__ __forceinline__ void block_coords(int64_t nbm, int64_t nbn,
                                             int64_t *bm, int64_t *bn) {
  const int64_t total = nbm * nbn;
  const int64_t L = blockIdx.x;
  int64_t lin = L;
  if (total % 8 == 0) lin = (L % 8) * (total / 8) + L / 8;
  const int64_t G = 8;
  const int64_t group = lin / (G * nbn);
  const int64_t first = group * G;
  const int64_t gsize = (nbm - first) < G ? (nbm - first) : G;
  const int64_t in = lin - group * G * nbn;
  *bm = first + in % gsize;
  *bn = in / gsize;
}

__device__ __forceinline__ void store_c(const DenseParams &p, int64_t row,
                                        int64_t col, float acc) {
  if (row < p.M && col < p.N) {
    float *cp = p.C + row * p.sCm + col * p.sCn;
    float v = p.alpha * acc;
    if (p.read_c) v = p.alpha * acc + p.beta * (*cp);
    *cp = v;
  }
}

// ---- split-precision bf16 GEMM ---------------------------------------------
// bf16 element offset of 16-byte chunk kq (0..3) of LDS tile row x (32 bf16
// = 64 bytes per row, unpadded).  The chunk is XORed with
// f(x) = ((x >> 1) ^ (x >> 2)) & 3, which takes 4 distinct values both over
// rows of one parity among 8 consecutive rows and over rows of one residue
// mod 4 among 16: 8 (16) consecutive rows reading one chunk then hit 8 (16)
// distinct 16-byte bank groups, whether the LDS serves 128 or 256 bytes per
// clock, and 4 threads writing one row's 4 chunks stay contiguous.  (With
// f(x) = (x >> 2) & 3, rows x and x + 2 collide in the 128-byte case.)
__device__ __forceinline__ int sw_off(int x, int kq) {
  return x * kSplitKBlock + 8 * (kq ^ (((x >> 1) ^ (x >> 2)) & 3));
}

template <int S>
__global__ __launch_bounds__(512) void k_dense_bf16(DenseParams p) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][S][BMH * kSplitKBlock];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][S][BNH * kSplitKBlock];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;  // 4 x 2 waves of 64 x 64
  int64_t bm, bn;
  block_coords(p.Mp / BMH, p.Np / BNH, &bm, &bn);
  const int64_t nsteps = p.Kp / kSplitKBlock;
  const uint16_t *Ag = (const uint16_t *)p.Ap + bm * BMH * kSplitKBlock;
  const uint16_t *Bg = (const uint16_t *)p.Bp + bn * BNH * kSplitKBlock;
  const int64_t planeA = p.Mp * p.Kp, planeB = p.Np * p.Kp;
  const int64_t slabA = p.Mp * kSplitKBlock, slabB = p.Np * kSplitKBlock;

  // Step t of the block: A rows bm*256 .. +255 (thread chunks c = tid and
  // tid + 512: row c / 4, chunk c % 4) and B rows bn*128 .. +127 (c = tid),
  // each a contiguous run of the k-blocked planes.
  u4 ra[S][2], rb[S];
  auto gload = [&](int64_t t) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
        ra[s][i] = *reinterpret_cast<const u4 *>(Ag + s * planeA + t * slabA + (tid + 512 * i) * 8);
      rb[s] = *reinterpret_cast<const u4 *>(Bg + s * planeB + t * slabB + tid * 8);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int c = tid + 512 * i;
        *reinterpret_cast<u4 *>(&As[buf][s][sw_off(c >> 2, c & 3)]) = ra[s][i];
      }
      *reinterpret_cast<u4 *>(&Bs[buf][s][sw_off(tid >> 2, tid & 3)]) = rb[s];
    }
  };

  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // Step t: read the fragments of step t from LDS buffer t % 2, store step
  // t + 1 (in registers since step t - 1) into buffer (t + 1) % 2 -- read
  // by no wave since the barrier that ended step t - 1 -- and reload the
  // registers with step t + 2; the LDS writes and the global loads then
  // run under the MFMAs of step t.  One barrier per step.
  const int r16 = lane & 15, g = lane >> 4;
  gload(0);
  lstore(0);
  if (nsteps > 1) gload(1);
  __syncthreads();
  // planes in the order the products below first need them
  constexpr int RA3[3] = {1, 2, 0}, RB3[3] = {1, 0, 2};
  constexpr int RA2[2] = {1, 0}, RB2[2] = {0, 1};
  for (int64_t t = 0; t < nsteps; ++t) {
    const int buf = (int)(t & 1);
    bf8 fa[S][4], fb[S][4];
#pragma unroll
    for (int q = 0; q < S; ++q) {
      const int sa = S == 3 ? RA3[q] : RA2[q];
      const int sb = S == 3 ? RB3[q] : RB2[q];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        fa[sa][i] = *reinterpret_cast<const bf8 *>(&As[buf][sa][sw_off(wr * 64 + i * 16 + r16, g)]);
        fb[sb][i] = *reinterpret_cast<const bf8 *>(&Bs[buf][sb][sw_off(wc * 64 + i * 16 + r16, g)]);
      }
    }
    if (t + 1 < nsteps) {
      lstore(buf ^ 1);
      if (t + 2 < nsteps) gload(t + 2);
    }
    // products, smallest terms first: (S==3) m*m, l*h, h*l, m*h, h*m, h*h;
    // (S==2) l*h, h*l, h*h.  The product loop is outermost so that
    // consecutive MFMAs update different tiles (no back-to-back dependency);
    // each tile still accumulates its products in the order above.
    constexpr int NP = S == 3 ? 6 : 3;
    constexpr int PA3[6] = {1, 2, 0, 1, 0, 0}, PB3[6] = {1, 0, 2, 0, 1, 0};
    constexpr int PA2[3] = {1, 0, 0}, PB2[3] = {0, 1, 0};
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const int ia = S == 3 ? PA3[q] : PA2[q];
      const int ib = S == 3 ? PB3[q] : PB2[q];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ia][i], fb[ib][j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        store_c(p, bm * BMH + wr * 64 + i * 16 + 4 * g + e,
                bn * BNH + wc * 64 + j * 16 + r16, acc[i][j][e]);
}

// ---- plain FP32 GEMM on v_mfma_f32_16x16x4f32 ------------------------------
__global__ __launch_bounds__(256) void k_dense_f32(DenseParams p) {
  __shared__ __attribute__((aligned(16))) float As[2][BM][BKF + PADF];
  __shared__ __attribute__((aligned(16))) float Bs[2][BN][BKF + PADF];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  int64_t bm, bn;
  block_coords(p.Mp / BM, p.Np / BN, &bm, &bn);
  const float *Ag = (const float *)p.Ap + bm * BM * p.Kp;
  const float *Bg = (const float *)p.Bp + bn * BN * p.Kp;
  f4 ra[2], rb[2];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = tid + 256 * i, row = c >> 2, kq = c & 3;
      ra[i] = *reinterpret_cast<const f4 *>(Ag + row * p.Kp + k0 + 4 * kq);
      rb[i] = *reinterpret_cast<const f4 *>(Bg + row * p.Kp + k0 + 4 * kq);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = tid + 256 * i, row = c >> 2, kq = c & 3;
      *reinterpret_cast<f4 *>(&As[buf][row][4 * kq]) = ra[i];
      *reinterpret_cast<f4 *>(&Bs[buf][row][4 * kq]) = rb[i];
    }
  };
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int r16 = lane & 15, g = lane >> 4;
  const int64_t nsteps = p.Kp / BKF;
  gload(0);
  lstore(0);
  __syncthreads();
  for (int64_t t = 0; t < nsteps; ++t) {
    const int buf = (int)(t & 1);
    if (t + 1 < nsteps) gload((t + 1) * BKF);
    f4 fa[4], fb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      fa[i] = *reinterpret_cast<const f4 *>(&As[buf][wr * 64 + i * 16 + r16][4 * g]);
      fb[i] = *reinterpret_cast<const f4 *>(&Bs[buf][wc * 64 + i * 16 + r16][4 * g]);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[i][e], fb[j][e], acc[i][j], 0, 0, 0);
    if (t + 1 < nsteps) lstore(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        store_c(p, bm * BM + wr * 64 + i * 16 + 4 * g + e,
                bn * BN + wc * 64 + j * 16 + r16, acc[i][j][e]);
}

}  // namespace

hipError_t launch_split(const float *src, int64_t sR, int64_t sK, int64_t rows,
                        int64_t K, int64_t rowsP, int64_t Kp, int S, void *dst,
                        hipStream_t st) {
  const int64_t total = rowsP * (Kp / 8);
  if (total <= 0) return hipSuccess;
  const int kfast = sK == 1 ? 1 : 0;
  const dim3 grid((unsigned)((total + 255) / 256));
  switch (S) {
    case 0: hipLaunchKernelGGL(k_split<0>, grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, dst, kfast); break;
    case 2: hipLaunchKernelGGL(k_split<2>, grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, dst, kfast); break;
    case 3: hipLaunchKernelGGL(k_split<3>, grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, dst, kfast); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_dense(int S, const DenseParams &p, hipStream_t st) {
  const int64_t nblk = S == 0 ? (p.Mp / BM) * (p.Np / BN) : (p.Mp / BMH) * (p.Np / BNH);
  if (nblk <= 0) return hipSuccess;
  if (nblk > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)nblk);
  switch (S) {
    case 0: hipLaunchKernelGGL(k_dense_f32, grid, dim3(256), 0, st, p); break;
    case 2: hipLaunchKernelGGL(k_dense_bf16<2>, grid, dim3(512), 0, st, p); break;
    case 3: hipLaunchKernelGGL(k_dense_bf16<3>, grid, dim3(512), 0, st, p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hml
