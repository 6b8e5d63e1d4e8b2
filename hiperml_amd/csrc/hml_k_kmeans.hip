// hml_k_kmeans.hip -- k-means nearest-centroid assignment on gfx950.
//
// Replaces the reference's assign kernels hmlKmeansAssignCtrdSmem /
// SmemUnroll / Gmem (include/hml_kmeans_kernel_template.h:18-88, 90-316,
// 323-367), which give each thread one row and loop over every centroid
// with scalar FMAs.  Their result is defined by that arithmetic: per
// centroid k ascending, dist = sum over dims ascending of delta*delta
// (delta = x - c, one rounding per step), first k strictly below the
// running minimum (from FLT_MAX).  Same result here, different machine:
//
//  1. Screening on the FP32 matrix cores.  |x - c|^2 = |x|^2 + s_k with
//     s_k = |c_k|^2 - 2 x.c_k, and x.c_k for a 16-row tile against 16
//     centroids is one v_mfma_f32_16x16x4f32 per 4 dims: a skinny GEMM
//     (rows x dims) * (dims x centroids).  Lane l = 16q + r holds row r,
//     dims 16i + 4q .. +3 (i < NI) from row-contiguous loads; the centroid
//     operand uses the same dim order, so MFMA j = 4i + e, k-slot q covers
//     dim 16i + 4q + e.  The result lane l holds x_r . c_{16t + 4q + i}.
//  2. A bound on how far the reference's float distance can be from
//     |x|^2 + s~_k (the screened value): with S_k = (|x| + |c_k|)^2 +
//     |s~_k|, the reference's rounding (D + 3 roundings of nonnegative
//     terms) and the screen's (D-term dot and norm, one subtraction) are
//     each below (2D + 4) 2^-24 S_k; the margin m_k = 4 (D + 4) 2^-24 S_k
//     covers both with slack for rounding the margin itself.  Only
//     centroids with s~_k - m_k <= min_j (s~_j + m_j) can be the exact
//     argmin; NaN or infinite screens make every centroid a candidate.
//  3. Exact refinement of the candidates (usually one per row) with the
//     reference's own arithmetic, from LDS copies of the row and the
//     centroid; the four lanes of a row refine disjoint candidate sets in
//     ascending k and combine by (distance, index), which is the
//     reference's "first strict minimum".
//
// Centroids live in LDS (zero padded to 16 NI dims, row stride 16 NI + 4
// floats), staged once per workgroup when they fit, else chunk by chunk per
// row step with the workgroup's waves in lockstep.  Candidates are found
// per pass of NT x 16 centroids; passes and chunks combine in ascending k
// with a strict comparison, so any partition gives the same answer.
#include <float.h>

#include "hml_internal.h"
#include "hml_kdev.h"

namespace hml {
namespace {

constexpr uint32_t kNone = 0xffffffffu;

struct KmParams {
  const float *rows;
  const float *ctrds;
  uint32_t *assign;
  int64_t nrows;
  int64_t steps;  // row steps of kKmWaves tiles per workgroup (uniform)
  int32_t D, K;
  int32_t KC;     // centroids per LDS chunk (multiple of 16 NT)
};

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exact distance of LDS row xr to LDS centroid cr over nd4 groups of 4
// dims (dims past D are zero in both, and adding (+0)^2 to a nonnegative
// sum changes nothing).
template <bool FUSED>
__device__ __forceinline__ float exact_dist(const float *xr, const float *cr, int nd4) {
  float dist = 0.f;
  for (int d = 0; d < nd4; ++d) {
    const f4 xv = *reinterpret_cast<const f4 *>(xr + 4 * d);
    const f4 cv = *reinterpret_cast<const f4 *>(cr + 4 * d);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float delta = xv[e] - cv[e];
      if (FUSED) {
        dist = __builtin_fmaf(delta, delta, dist);
      } else {
#pragma clang fp contract(off)
        const float sq = delta * delta;
        dist = dist + sq;
      }
    }
  }
  return dist;
}

template <int NI, int NT, bool FUSED>
__global__ __launch_bounds__(64 * kKmWaves) void k_kmeans_assign(KmParams p) {
  constexpr int DP = 16 * NI + 4;  // LDS row stride (floats), 16-byte multiple
  extern __shared__ f4 km_lds4[];
  float *lds = reinterpret_cast<float *>(km_lds4);
  float *cs = lds;                   // KC x DP centroids of the current chunk
  float *cn = cs + p.KC * DP;        // their squared norms
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float *xs = cn + p.KC + w * 16 * DP;  // this wave's 16 rows
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, r = lane & 15;
  const int D = p.D, K = p.K;
  const int nd4 = (D + 3) >> 2;
  const int nchunks = (K + p.KC - 1) / p.KC;
  const float tau = 4.0f * (float)(D + 4) * 0x1p-24f;

  auto stage = [&](int c0) {
    __syncthreads();  // every wave is done with the previous chunk
    for (int idx = threadIdx.x; idx < p.KC * DP; idx += blockDim.x) {
      const int k = idx / DP, d = idx - k * DP;
      cs[idx] = (c0 + k < K && d < D) ? p.ctrds[(int64_t)(c0 + k) * D + d] : 0.f;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < p.KC; k += blockDim.x) {
      float s = 0.f;
      for (int d = 0; d < D; ++d) s = __builtin_fmaf(cs[k * DP + d], cs[k * DP + d], s);
      cn[k] = s;
    }
    __syncthreads();
  };
  if (nchunks == 1) stage(0);

  for (int64_t st = 0; st < p.steps; ++st) {
    const int64_t tile = (st * gridDim.x + blockIdx.x) * kKmWaves + w;
    const int64_t row0 = tile * 16;
    const bool live = row0 < p.nrows;  // wave-uniform
    if (nchunks == 1 && !live) break;  // later steps only go further

    f4 xv[NI];
    float xx = 0.f;
    if (live) {
      const int64_t left = (p.nrows - row0) * D * 4;
      const __amdgpu_buffer_rsrc_t rs =
          make_rsrc(p.rows + row0 * D, left < kRsrcMax ? (int32_t)left : kRsrcMax);
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        const int d0 = 16 * i + 4 * q;
        const bool rin = row0 + r < p.nrows;
        const uint32_t o = (uint32_t)((r * D + d0) * 4);
        if (d0 + 4 <= D) {
          xv[i] = buf_load4<kBufNT>(rs, rin ? o : kOOB, 0);
        } else {
          // a group that crosses the end of the row, element by element
          // (its tail belongs to the next row, or lies past the array)
#pragma unroll
          for (int e = 0; e < 4; ++e)
            xv[i][e] = buf_load1(rs, rin && d0 + e < D ? o + 4 * e : kOOB, 0);
        }
      }
      wave_lds_sync();  // the previous tile's refinement reads are done
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        *reinterpret_cast<f4 *>(xs + r * DP + 16 * i + 4 * q) = xv[i];
#pragma unroll
        for (int e = 0; e < 4; ++e) xx = __builtin_fmaf(xv[i][e], xv[i][e], xx);
      }
      xx += __shfl_xor(xx, 16);
      xx += __shfl_xor(xx, 32);
      wave_lds_sync();
    }

    float rbd = FLT_MAX;
    uint32_t rbk = kNone;
    for (int c = 0; c < nchunks; ++c) {
      if (nchunks > 1) stage(c * p.KC);
      if (!live) continue;
      const int kc = min(p.KC, K - c * p.KC);
      for (int pb = 0; pb < kc; pb += 16 * NT) {
        f4 acc[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int i = 0; i < NI; ++i) {
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const f4 cf = *reinterpret_cast<const f4 *>(cs + (pb + 16 * t + r) * DP + 16 * i + 4 * q);
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[t] = mfma4(cf[e], xv[i][e], acc[t]);
          }
        }
        // screen: threshold over the pass, then candidates
        float T = __builtin_inff();
        f4 nrm[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          nrm[t] = *reinterpret_cast<const f4 *>(cn + pb + 16 * t + 4 * q);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float s = nrm[t][i] - 2.f * acc[t][i];
            const float m = tau * (xx + nrm[t][i] + 2.f * __builtin_sqrtf(xx * nrm[t][i]) + __builtin_fabsf(s));
            if (pb + 16 * t + 4 * q + i < kc) T = __builtin_fminf(T, s + m);
          }
        }
        T = __builtin_fminf(T, __shfl_xor(T, 16));
        T = __builtin_fminf(T, __shfl_xor(T, 32));
        uint32_t cand = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const float s = nrm[t][i] - 2.f * acc[t][i];
            const float m = tau * (xx + nrm[t][i] + 2.f * __builtin_sqrtf(xx * nrm[t][i]) + __builtin_fabsf(s));
            if (pb + 16 * t + 4 * q + i < kc && !(s - m > T)) cand |= 1u << (4 * t + i);
          }
        // exact refinement, ascending k within the lane
        float bd = FLT_MAX;
        uint32_t bk = kNone;
        while (cand) {
          const int b = __builtin_ctz(cand);
          cand &= cand - 1;
          const int kl = pb + 16 * (b >> 2) + 4 * q + (b & 3);
          const float dist = exact_dist<FUSED>(xs + r * DP, cs + kl * DP, nd4);
          if (dist < bd) {
            bd = dist;
            bk = (uint32_t)(c * p.KC + kl);
          }
        }
#pragma unroll
        for (int off = 16; off < 64; off <<= 1) {
          const float obd = __shfl_xor(bd, off);
          const uint32_t obk = (uint32_t)__shfl_xor((int)bk, off);
          if (obd < bd || (obd == bd && obk < bk)) {
            bd = obd;
            bk = obk;
          }
        }
        if (bd < rbd) {  // earlier passes hold smaller k: they keep ties
          rbd = bd;
          rbk = bk;
        }
      }
    }
    if (live && q == 0 && row0 + r < p.nrows)
      p.assign[row0 + r] = rbk == kNone ? (uint32_t)K : rbk;
  }
}

typedef void (*KmKernel)(KmParams);

template <int NI, int NT>
KmKernel km_pick(bool fused) {
  return fused ? k_kmeans_assign<NI, NT, true> : k_kmeans_assign<NI, NT, false>;
}

template <int NI>
KmKernel km_pick_nt(int nt, bool fused) {
  switch (nt) {
    case 1: return km_pick<NI, 1>(fused);
    case 2: return km_pick<NI, 2>(fused);
    case 4: return km_pick<NI, 4>(fused);
    default: return km_pick<NI, 8>(fused);
  }
}

KmKernel km_kernel(int ni, int nt, bool fused) {
  switch (ni) {
    case 1: return km_pick_nt<1>(nt, fused);
    case 2: return km_pick_nt<2>(nt, fused);
    case 3: return km_pick_nt<3>(nt, fused);
    default: return km_pick_nt<4>(nt, fused);
  }
}

int cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    return 256;
  return n;
}

}  // namespace

KmeansPlan plan_kmeans(int64_t nrows, int D, int K) {
  KmeansPlan pl;
  pl.ni = (D + 15) / 16;
  pl.nt = K <= 16 ? 1 : K <= 32 ? 2 : K <= 64 ? 4 : 8;
  const int dp = 16 * pl.ni + 4;
  const int pass = 16 * pl.nt;
  // LDS: centroid chunk (KC x dp) + norms (KC) + 16 rows per wave, at most
  // kKmLdsBytes.
  const int64_t rows_f = (int64_t)kKmWaves * 16 * dp;
  const int64_t room = kKmLdsBytes / 4 - rows_f;
  int64_t kc = room / (dp + 1) / pass * pass;
  const int64_t kpad = ((int64_t)K + pass - 1) / pass * pass;
  if (kc > kpad) kc = kpad;
  pl.KC = (int)kc;
  pl.lds = (size_t)(kc * (dp + 1) + rows_f) * 4;
  pl.tiles = (nrows + 15) / 16;
  return pl;
}

hipError_t launch_kmeans_assign(const KmeansPlan &pl, uint32_t *assign, const float *rows,
                                int64_t nrows, int D, const float *ctrds, int K, bool fused,
                                hipStream_t st) {
  if (pl.KC < 16 * pl.nt || nrows <= 0) return nrows <= 0 ? hipSuccess : hipErrorInvalidValue;
  KmKernel fn = km_kernel(pl.ni, pl.nt, fused);
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(fn),
                                                              64 * kKmWaves, pl.lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const int64_t wgs_needed = (pl.tiles + kKmWaves - 1) / kKmWaves;
  int64_t grid = (int64_t)per_cu * cu_count();
  if (grid > wgs_needed) grid = wgs_needed;
  KmParams p;
  p.rows = rows;
  p.ctrds = ctrds;
  p.assign = assign;
  p.nrows = nrows;
  p.steps = (pl.tiles + grid * kKmWaves - 1) / (grid * kKmWaves);
  p.D = D;
  p.K = K;
  p.KC = pl.KC;
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64 * kKmWaves), pl.lds, st, p);
  return hipGetLastError();
}

}  // namespace hml
