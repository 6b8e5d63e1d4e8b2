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
// Scaled fp16 split ("f16x3"; Ootomo & Yokota, "Recovering single precision
// accuracy from Tensor Cores while surpassing the FP32 theoretical peak
// performance", IJHPCA 2022): the same three products h*h + h*l + l*h, but
// with fp16 terms, whose 11-bit significands make the two-term split 22 bits
// wide (bf16: 16), so the representation error per operand is ~2^-22
// instead of ~2^-16.  fp16's narrow exponent range is handled by exact
// power-of-two scaling: row r of A (column n of B) is multiplied by 2^e_r
// with e_r chosen so that its largest magnitude lands in [2^14, 2^15), and
// the epilogue multiplies by 2^-(e_m + e_n).  The correction products
// h*l + l*h accumulate separately from h*h (their roundings then act on a
// 2^-11 smaller sum) and the two sums meet once, in the epilogue.
// v_mfma_f32_16x16x32_f16 runs at the bf16 rate.
//
// Data path: a split pass writes zero-padded 16-bit planes in a k-blocked
// layout, Ap[S][Kp/32][Mp][32] and Bp[S][Kp/32][Np][32] (B transposed): the
// 32-wide k-step of a 256-row panel is then 16 KB of contiguous memory, read
// with whole-line 16-byte loads and no bounds checks.  Block tile 256 x 128 x
// 32 (128 x 128 on short K), 8 (4) waves in a 4 x 2 (2 x 2) grid of 64 x 64
// wave tiles (2 waves per SIMD), two
// LDS buffers (one barrier per k-step), global loads of step t+1 in flight
// while step t computes, LDS rows unpadded with the 16-byte chunk index
// XOR-swizzled by the row (conflict-free row-contiguous writes and operand
// reads), and an XCD-aware block order (blocks that share A/B panels run on
// the same L2).  The plain FP32 kernel keeps K-contiguous FP32 copies and
// 128 x 128 tiles.
#include <algorithm>

#include "hml_internal.h"
#include "hml_dense.h"

namespace hml {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int BM = kDenseTileM, BN = kDenseTileN;  // FP32 kernel tile
constexpr int BNH = kSplitTileN;                    // 16-bit kernels' tile width
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

// ---- FP32 pack pass -----------------------------------------------------------
// K-contiguous FP32 copy [rowsP][Kp] of a rows x K operand X(r,k) =
// src[r*sR + k*sK], zero-padded, for the plain FP32 kernel.  One thread per
// 8 consecutive k of one row; the thread index runs fastest along whichever
// of r / k is contiguous in src, so the reads coalesce.
__global__ void k_pack_f32(const float *__restrict__ src, int64_t sR, int64_t sK,
                           int64_t rows, int64_t K, int64_t rowsP, int64_t Kp,
                           float *__restrict__ dst, int kfast) {
  const int64_t kc_n = Kp / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rowsP * kc_n) return;
  int64_t r, kc;
  if (kfast) {
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
  float *d = dst + r * Kp + kc * 8;
  *reinterpret_cast<f4 *>(d) = f4{v[0], v[1], v[2], v[3]};
  *reinterpret_cast<f4 *>(d + 4) = f4{v[4], v[5], v[6], v[7]};
}

// ---- tiled split passes ------------------------------------------------------
// The bf16 / fp16 split passes and the row-maximum pass walk the operand in
// tiles of 64 rows x 32 k (one k-block of the plane layout).  A tile is read
// with the operand's own contiguous dimension across the lanes -- 16-byte
// loads, 128 contiguous bytes per row when K is contiguous (LM_K), 256 per
// k when rows are (LM_R), element loads otherwise (LM_ANY) -- into an LDS
// tile (row stride 33 floats), then thread t takes row t / 4, k-chunk
// t % 4 (8 consecutive k): the split writes that chunk's 16 bytes of each
// plane, and the 256 threads of a block write 64 rows x 64 bytes = 4 KB of
// every plane contiguously.
constexpr int kLmK = 0, kLmR = 1, kLmAny = 2;
constexpr int kSplitRows = 64;
constexpr int kLdsStride = 33;

// X(r0 + i, k0 + j), i < 64, j < 32, into lt[i * 33 + j]; zero outside the
// operand.  vec: 16-byte loads are aligned (host-checked).
template <int LM>
__device__ __forceinline__ void load_split_tile(const float *__restrict__ src, int64_t sR,
                                                int64_t sK, int64_t rows, int64_t K,
                                                int64_t r0, int64_t k0, int vec, float *lt) {
  const int t = threadIdx.x;
  if constexpr (LM == kLmK) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int i = (t >> 3) + 32 * it, j = (t & 7) * 4;
      const int64_t r = r0 + i, k = k0 + j;
      f4 v = {0.f, 0.f, 0.f, 0.f};
      if (r < rows) {
        const float *p = src + r * sR + k;
        if (vec && k + 3 < K) {
          v = *reinterpret_cast<const f4 *>(p);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = k + e < K ? p[e] : 0.f;
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) lt[i * kLdsStride + j + e] = v[e];
    }
  } else if constexpr (LM == kLmR) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      const int j = (t >> 4) + 16 * it, i = (t & 15) * 4;
      const int64_t r = r0 + i, k = k0 + j;
      f4 v = {0.f, 0.f, 0.f, 0.f};
      if (k < K) {
        const float *p = src + k * sK + r;
        if (vec && r + 3 < rows) {
          v = *reinterpret_cast<const f4 *>(p);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = r + e < rows ? p[e] : 0.f;
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) lt[(i + e) * kLdsStride + j] = v[e];
    }
  } else {
#pragma unroll
    for (int it = 0; it < 8; ++it) {
      const int idx = t + 256 * it, i = idx >> 5, j = idx & 31;
      const int64_t r = r0 + i, k = k0 + j;
      lt[i * kLdsStride + j] = (r < rows && k < K) ? src[r * sR + k * sK] : 0.f;
    }
  }
}

// rowmax[r] = max_k |X(r,k)| as FP32 bits (non-negative floats order like
// their bit patterns, so an integer atomicMax is a float max; NaN ranks
// above Inf).  rowmax must be zero on entry.  Block b covers rows
// 64 (b % nrb) .. +63 and k-blocks kpb (b / nrb) .. +kpb-1 (nrb row
// blocks); one atomic per row and block.
template <int LM>
__global__ __launch_bounds__(256) void k_rowmax(const float *__restrict__ src, int64_t sR,
                                                int64_t sK, int64_t rows, int64_t K, int vec,
                                                int kpb, uint32_t *__restrict__ rowmax) {
  __shared__ float lt[kSplitRows * kLdsStride];
  const int t = threadIdx.x, i = t >> 2, q = t & 3;
  const int64_t nrb = (rows + kSplitRows - 1) / kSplitRows;
  const int64_t r0 = ((int64_t)blockIdx.x % nrb) * kSplitRows;
  const int64_t kt0 = ((int64_t)blockIdx.x / nrb) * kpb;
  const int64_t nkt = (K + kSplitKBlock - 1) / kSplitKBlock;
  uint32_t m = 0;
  for (int64_t kt = kt0; kt < kt0 + kpb && kt < nkt; ++kt) {
    load_split_tile<LM>(src, sR, sK, rows, K, r0, kt * kSplitKBlock, vec, lt);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j)
      m = max(m, __float_as_uint(lt[i * kLdsStride + q * 8 + j]) & 0x7fffffffu);
    __syncthreads();
  }
  // the 4 threads of a row are adjacent lanes
  m = max(m, (uint32_t)__shfl_xor((int)m, 1, 64));
  m = max(m, (uint32_t)__shfl_xor((int)m, 2, 64));
  if (q == 0 && m && r0 + i < rows) atomicMax(rowmax + r0 + i, m);
}

// Exponent e with max * 2^e in [2^14, 2^15) (0 for a zero, Inf or NaN max).
__device__ __forceinline__ int split_exponent(uint32_t maxbits) {
  const float mx = __uint_as_float(maxbits);
  if (maxbits == 0 || maxbits >= 0x7f800000u) return 0;
  return 14 - ilogbf(mx);
}

// bf16 planes (S = 2: hi, lo; S = 3: hi, mid, lo; each residual exact in
// FP32, rounding to nearest even) or, F16, scaled fp16 hi / lo planes
// (row r times 2^e_r from rowmax, then hi = fp16(x), lo = fp16(x - hi)),
// in the k-blocked layout dst[s][Kp/32][rowsP][32].
template <int S, bool F16, int LM>
__global__ __launch_bounds__(256) void k_split_tiled(const float *__restrict__ src, int64_t sR,
                                                     int64_t sK, int64_t rows, int64_t K,
                                                     int64_t rowsP, int64_t Kp, int vec,
                                                     const uint32_t *__restrict__ rowmax,
                                                     void *__restrict__ dst) {
  __shared__ float lt[kSplitRows * kLdsStride];
  const int t = threadIdx.x, i = t >> 2, q = t & 3;
  // block b: row block b % nrb of k-block b / nrb, so consecutive blocks
  // write consecutive 4 KB pieces of each plane
  const int64_t nrb = rowsP / kSplitRows;
  const int64_t r0 = ((int64_t)blockIdx.x % nrb) * kSplitRows;
  const int64_t kt = (int64_t)blockIdx.x / nrb;
  load_split_tile<LM>(src, sR, sK, rows, K, r0, kt * kSplitKBlock, vec, lt);
  __syncthreads();
  float v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = lt[i * kLdsStride + q * 8 + j];
  const int64_t plane = rowsP * Kp;
  const int64_t off = (kt * rowsP + r0 + i) * kSplitKBlock + q * 8;
  if constexpr (F16) {
    const int e = r0 + i < rows ? split_exponent(rowmax[r0 + i]) : 0;
    _Float16 hi[8], lo[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = ldexpf(v[j], e);
      hi[j] = (_Float16)x;
      lo[j] = (_Float16)(x - (float)hi[j]);  // the residual is exact in FP32
    }
    _Float16 *d = (_Float16 *)dst + off;
    *reinterpret_cast<h8 *>(d) = h8{hi[0], hi[1], hi[2], hi[3], hi[4], hi[5], hi[6], hi[7]};
    *reinterpret_cast<h8 *>(d + plane) = h8{lo[0], lo[1], lo[2], lo[3], lo[4], lo[5], lo[6], lo[7]};
  } else {
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
      *reinterpret_cast<u4 *>((uint16_t *)dst + s * plane + off) = w;
    }
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

// S = 2 / 3: bf16 planes, one accumulator; S = 2 with F16: scaled fp16
// planes, h*h and the corrections in separate accumulators.  WM waves along
// M: a block tile of 64 WM x 128 (WM = 4: 256 x 128, 8 waves, one workgroup
// per CU by LDS; WM = 2: 128 x 128, 4 waves, two workgroups per CU, so one
// workgroup's C stores run under the other's products on short K).
template <int S, bool F16, int WM>
__global__ __launch_bounds__(128 * WM, 2) void k_dense_bf16(DenseParams p) {
  constexpr int TM = 64 * WM;       // block rows
  constexpr int NT = 128 * WM;      // threads
  constexpr int NA = TM * 4 / NT;   // 16-byte A chunks per thread and step
  constexpr int NB = BNH * 4 / NT;  // 16-byte B chunks per thread and step
  __shared__ __attribute__((aligned(16))) uint16_t As[2][S][TM * kSplitKBlock];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[2][S][BNH * kSplitKBlock];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;  // WM x 2 waves of 64 x 64
  int64_t bm, bn;
  block_coords(p.Mp / TM, p.Np / BNH, &bm, &bn);
  const int64_t nsteps = p.Kp / kSplitKBlock;
  const uint16_t *Ag = (const uint16_t *)p.Ap + bm * TM * kSplitKBlock;
  const uint16_t *Bg = (const uint16_t *)p.Bp + bn * BNH * kSplitKBlock;
  const int64_t planeA = p.Mp * p.Kp, planeB = p.Np * p.Kp;
  const int64_t slabA = p.Mp * kSplitKBlock, slabB = p.Np * kSplitKBlock;

  // Step t of the block: A rows bm*TM .. +TM-1 (thread chunks c = tid +
  // NT i: row c / 4, chunk c % 4) and B rows bn*128 .. +127 (likewise), each
  // a contiguous run of the k-blocked planes.
  u4 ra[S][NA], rb[S][NB];
  auto gload = [&](int64_t t) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < NA; ++i)
        ra[s][i] = *reinterpret_cast<const u4 *>(Ag + s * planeA + t * slabA + (tid + NT * i) * 8);
#pragma unroll
      for (int i = 0; i < NB; ++i)
        rb[s][i] = *reinterpret_cast<const u4 *>(Bg + s * planeB + t * slabB + (tid + NT * i) * 8);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
#pragma unroll
      for (int i = 0; i < NA; ++i) {
        const int c = tid + NT * i;
        *reinterpret_cast<u4 *>(&As[buf][s][sw_off(c >> 2, c & 3)]) = ra[s][i];
      }
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int c = tid + NT * i;
        *reinterpret_cast<u4 *>(&Bs[buf][s][sw_off(c >> 2, c & 3)]) = rb[s][i];
      }
    }
  };

  f4 acc[4][4], acc2[F16 ? 4 : 1][F16 ? 4 : 1];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  if (F16) {
#pragma unroll
    for (int i = 0; i < (F16 ? 4 : 1); ++i)
#pragma unroll
      for (int j = 0; j < (F16 ? 4 : 1); ++j) acc2[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  }

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
        for (int j = 0; j < 4; ++j) {
          if constexpr (F16) {
            // q = 0, 1: the corrections l*h, h*l; q = 2: h*h
            const h8 a = __builtin_bit_cast(h8, fa[ia][i]);
            const h8 b = __builtin_bit_cast(h8, fb[ib][j]);
            if (q < 2) acc2[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc2[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[i][j], 0, 0, 0);
          } else {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[ia][i], fb[ib][j], acc[i][j], 0, 0, 0);
          }
        }
    }
    __syncthreads();
  }

  if constexpr (F16) {
    // undo the row / column scaling: exact powers of two
    int ea[4][4], eb[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) eb[j] = split_exponent(p.rmaxB[bn * BNH + wc * 64 + j * 16 + r16]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int e = 0; e < 4; ++e) ea[i][e] = split_exponent(p.rmaxA[bm * TM + wr * 64 + i * 16 + 4 * g + e]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          store_c(p, bm * TM + wr * 64 + i * 16 + 4 * g + e,
                  bn * BNH + wc * 64 + j * 16 + r16,
                  ldexpf(acc[i][j][e] + acc2[i][j][e], -(ea[i][e] + eb[j])));
  } else {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          store_c(p, bm * TM + wr * 64 + i * 16 + 4 * g + e,
                  bn * BNH + wc * 64 + j * 16 + r16, acc[i][j][e]);
  }
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

namespace {

// Load mode of a rows x K operand with strides (sR, sK), and whether its
// 16-byte loads are aligned.
int split_load_mode(const float *src, int64_t sR, int64_t sK, int64_t rows, int64_t K,
                    int *vec) {
  const bool al = (uintptr_t)src % 16 == 0;
  if (sK == 1 || K == 1) {
    *vec = al && (sR % 4 == 0 || rows == 1);
    return kLmK;
  }
  if (sR == 1 || rows == 1) {
    *vec = al && sK % 4 == 0;
    return kLmR;
  }
  *vec = 0;
  return kLmAny;
}

template <int S, bool F16>
hipError_t launch_split_mode(int lm, dim3 grid, hipStream_t st, const float *src, int64_t sR,
                             int64_t sK, int64_t rows, int64_t K, int64_t rowsP, int64_t Kp,
                             int vec, const uint32_t *rowmax, void *dst) {
  switch (lm) {
    case kLmK: hipLaunchKernelGGL((k_split_tiled<S, F16, kLmK>), grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, vec, rowmax, dst); break;
    case kLmR: hipLaunchKernelGGL((k_split_tiled<S, F16, kLmR>), grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, vec, rowmax, dst); break;
    default: hipLaunchKernelGGL((k_split_tiled<S, F16, kLmAny>), grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp, vec, rowmax, dst); break;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_split(const float *src, int64_t sR, int64_t sK, int64_t rows,
                        int64_t K, int64_t rowsP, int64_t Kp, int S, void *dst,
                        hipStream_t st) {
  if (rowsP <= 0 || Kp <= 0) return hipSuccess;
  if (S == 0) {
    const int64_t total = rowsP * (Kp / 8);
    const int kfast = sK == 1 ? 1 : 0;
    const dim3 grid((unsigned)((total + 255) / 256));
    hipLaunchKernelGGL(k_pack_f32, grid, dim3(256), 0, st, src, sR, sK, rows, K, rowsP, Kp,
                       (float *)dst, kfast);
    return hipGetLastError();
  }
  const int64_t nblk = (rowsP / kSplitRows) * (Kp / kSplitKBlock);
  if (rowsP % kSplitRows != 0 || Kp % kSplitKBlock != 0 || nblk > 0x7fffffff)
    return hipErrorInvalidValue;
  int vec = 0;
  const int lm = split_load_mode(src, sR, sK, rows, K, &vec);
  const dim3 grid((unsigned)nblk);
  switch (S) {
    case 2: return launch_split_mode<2, false>(lm, grid, st, src, sR, sK, rows, K, rowsP, Kp, vec, nullptr, dst);
    case 3: return launch_split_mode<3, false>(lm, grid, st, src, sR, sK, rows, K, rowsP, Kp, vec, nullptr, dst);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_rowmax(const float *src, int64_t sR, int64_t sK, int64_t rows,
                         int64_t K, uint32_t *rowmax, hipStream_t st) {
  if (rows <= 0 || K <= 0) return hipSuccess;
  int vec = 0;
  const int lm = split_load_mode(src, sR, sK, rows, K, &vec);
  const int64_t nkt = (K + kSplitKBlock - 1) / kSplitKBlock;
  const int64_t nrb = (rows + kSplitRows - 1) / kSplitRows;
  // k-blocks per workgroup: up to 8 (fewer atomics), but enough workgroups
  // (about 4096) to fill the device when the operand is small
  const int kpb = (int)std::max<int64_t>(1, std::min<int64_t>(8, nrb * nkt / 4096));
  const int64_t nblk = nrb * ((nkt + kpb - 1) / kpb);
  if (nblk > 0x7fffffff) return hipErrorInvalidValue;
  const dim3 grid((unsigned)nblk);
  switch (lm) {
    case kLmK: hipLaunchKernelGGL(k_rowmax<kLmK>, grid, dim3(256), 0, st, src, sR, sK, rows, K, vec, kpb, rowmax); break;
    case kLmR: hipLaunchKernelGGL(k_rowmax<kLmR>, grid, dim3(256), 0, st, src, sR, sK, rows, K, vec, kpb, rowmax); break;
    default: hipLaunchKernelGGL(k_rowmax<kLmAny>, grid, dim3(256), 0, st, src, sR, sK, rows, K, vec, kpb, rowmax); break;
  }
  return hipGetLastError();
}

hipError_t launch_split_f16(const float *src, int64_t sR, int64_t sK, int64_t rows,
                            int64_t K, int64_t rowsP, int64_t Kp, const uint32_t *rowmax,
                            void *dst, hipStream_t st) {
  if (rowsP <= 0 || Kp <= 0) return hipSuccess;
  const int64_t nblk = (rowsP / kSplitRows) * (Kp / kSplitKBlock);
  if (rowsP % kSplitRows != 0 || Kp % kSplitKBlock != 0 || nblk > 0x7fffffff)
    return hipErrorInvalidValue;
  int vec = 0;
  const int lm = split_load_mode(src, sR, sK, rows, K, &vec);
  const dim3 grid((unsigned)nblk);
  return launch_split_mode<2, true>(lm, grid, st, src, sR, sK, rows, K, rowsP, Kp, vec, rowmax,
                                    dst);
}

// Block height of the two-plane kernels (bf16x3, f16x3): 128-row tiles
// (two workgroups per CU) when the k loop is short, 256-row tiles (more
// reuse per load) otherwise; HML_DENSE_TILE_M = 128 / 256 forces one (for
// measurements).  bf16x6's three planes leave LDS for one 128-row
// workgroup per CU only, so it always runs 256-row tiles.
int dense_tile_m(int S, int64_t Kp) {
  static const int force = [] {
    const char *e = getenv("HML_DENSE_TILE_M");
    return e ? atoi(e) : 0;
  }();
  if (S == 3) return 256;
  if (force == 128 || force == 256) return force;
  return Kp <= kDenseShortK ? 128 : 256;
}

hipError_t launch_dense(int S, const DenseParams &p, hipStream_t st) {
  // S: 0 FP32, 2 bf16x3, 3 bf16x6, 4 scaled f16x3
  const int tm = S == 0 ? BM : dense_tile_m(S, p.Kp);
  const int64_t nblk = S == 0 ? (p.Mp / BM) * (p.Np / BN) : (p.Mp / tm) * (p.Np / BNH);
  if (nblk <= 0) return hipSuccess;
  if (nblk > 0x7fffffff) return hipErrorInvalidConfiguration;
  const dim3 grid((unsigned)nblk);
  const bool t128 = tm == 128;
  switch (S) {
    case 0: hipLaunchKernelGGL(k_dense_f32, grid, dim3(256), 0, st, p); break;
    case 2:
      if (t128) hipLaunchKernelGGL((k_dense_bf16<2, false, 2>), grid, dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_dense_bf16<2, false, 4>), grid, dim3(512), 0, st, p);
      break;
    case 3: hipLaunchKernelGGL((k_dense_bf16<3, false, 4>), grid, dim3(512), 0, st, p); break;
    case 4:
      if (t128) hipLaunchKernelGGL((k_dense_bf16<2, true, 2>), grid, dim3(256), 0, st, p);
      else hipLaunchKernelGGL((k_dense_bf16<2, true, 4>), grid, dim3(512), 0, st, p);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace hml
