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

#include <utility>

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
  int64_t steps;  // row steps of kKmWaves wave tiles per workgroup (uniform)
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
// sum changes nothing).  The groups are read 4 at a time before their
// arithmetic, so the LDS latency is paid once per 16 dims.
template <bool FUSED>
__device__ __forceinline__ float exact_step(float dist, float x, float c) {
  const float delta = x - c;
  if (FUSED) return __builtin_fmaf(delta, delta, dist);
  {
#pragma clang fp contract(off)
    const float sq = delta * delta;
    return dist + sq;
  }
}

template <bool FUSED, int NI>
__device__ __forceinline__ float exact_dist(const float *xr, const float *cr, int nd4) {
  float dist = 0.f;
#pragma unroll
  for (int g = 0; g < NI; ++g) {
    if (4 * g < nd4) {
      f4 xv[4], cv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        xv[j] = *reinterpret_cast<const f4 *>(xr + 16 * g + 4 * j);
        cv[j] = *reinterpret_cast<const f4 *>(cr + 16 * g + 4 * j);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * g + j < nd4) {
#pragma unroll
          for (int e = 0; e < 4; ++e) dist = exact_step<FUSED>(dist, xv[j][e], cv[j][e]);
        }
    }
  }
  return dist;
}

// 128 centroids per pass over up to 32 dims: at most 128 VGPRs (4 waves per
// SIMD, two 8-wave workgroups per CU; without the bound the 32-dim tile took
// 129 and ran one workgroup per CU: 32 x 256 on 2^24 rows 3.81 -> 3.14 ms).
// The other tiles fit without it (bounding them measured slower), and the
// 33..64-dim ones would spill.
template <int NI, int NT, bool FUSED>
__global__ __launch_bounds__(64 * kKmWaves, (NT == 8 && NI <= 2) ? 4 : 1) void k_kmeans_assign(KmParams p) {
  constexpr int DP = 16 * NI + 4;  // LDS row stride (floats), 16-byte multiple
  extern __shared__ f4 km_lds4[];
  float *lds = reinterpret_cast<float *>(km_lds4);
  float *cs = lds;                   // KC x DP centroids of the current chunk
  float *cn = cs + p.KC * DP;        // their squared norms
  float *cmx = cn + p.KC;            // the largest norm of the chunk (4 floats)
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float *xs = cmx + 4 + w * 16 * DP;  // this wave's 16 rows
  const int lane = threadIdx.x & 63;
  const int q = lane >> 4, r = lane & 15;
  const int D = p.D, K = p.K;
  const int nd4 = (D + 3) >> 2;
  const int nchunks = (K + p.KC - 1) / p.KC;
  // margin per centroid: tau S_k (see the screen below); per row and chunk
  // the bound tau2 (|x| + max |c|)^2 >= tau S_k is used for every centroid
  const float tau = 4.0f * (float)(D + 4) * 0x1p-24f;
  const float tau2 = 2.0f * tau * (1.0f + 0x1p-10f);

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
    if (threadIdx.x < 64) {
      // largest squared norm (the zero padding adds nothing; a NaN norm is
      // skipped: that centroid's screen is NaN, which always passes)
      float mx = 0.f;
      for (int k = threadIdx.x; k < p.KC; k += 64) mx = __builtin_fmaxf(mx, cn[k]);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) mx = __builtin_fmaxf(mx, __shfl_xor(mx, off));
      if (threadIdx.x == 0) cmx[0] = __builtin_sqrtf(mx);
    }
    __syncthreads();
  };
  if (nchunks == 1) stage(0);

  // rows of step st's tile (zeros for a tile past the end); the next step's
  // rows are loaded while this step's are screened
  auto load_rows = [&](int64_t st, f4 (&dst)[NI]) {
    const int64_t row0 = ((st * gridDim.x + blockIdx.x) * kKmWaves + w) * 16;
    if (row0 >= p.nrows) {
#pragma unroll
      for (int i = 0; i < NI; ++i) dst[i] = f4{0.f, 0.f, 0.f, 0.f};
      return;
    }
    const int64_t left = (p.nrows - row0) * D * 4;
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(p.rows + row0 * D, left < kRsrcMax ? (int32_t)left : kRsrcMax);
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int d0 = 16 * i + 4 * q;
      const bool rin = row0 + r < p.nrows;
      const uint32_t o = (uint32_t)((r * D + d0) * 4);
      if (d0 + 4 <= D) {
        dst[i] = buf_load4<kBufNT>(rs, rin ? o : kOOB, 0);
      } else {
        // a group that crosses the end of the row, element by element
        // (its tail belongs to the next row, or lies past the array)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          dst[i][e] = buf_load1(rs, rin && d0 + e < D ? o + 4 * e : kOOB, 0);
      }
    }
  };
  f4 xnext[NI];
  load_rows(0, xnext);

  for (int64_t st = 0; st < p.steps; ++st) {
    const int64_t tile = (st * gridDim.x + blockIdx.x) * kKmWaves + w;
    const int64_t row0 = tile * 16;
    const bool live = row0 < p.nrows;  // wave-uniform
    if (nchunks == 1 && !live) break;  // later steps only go further

    f4 xv[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) xv[i] = xnext[i];
    if (st + 1 < p.steps) load_rows(st + 1, xnext);
    float xx = 0.f, xr = 0.f;
    if (live) {
      wave_lds_sync();  // the previous tile's refinement reads are done
#pragma unroll
      for (int i = 0; i < NI; ++i) {
        *reinterpret_cast<f4 *>(xs + r * DP + 16 * i + 4 * q) = xv[i];
#pragma unroll
        for (int e = 0; e < 4; ++e) xx = __builtin_fmaf(xv[i][e], xv[i][e], xx);
      }
      xx += __shfl_xor(xx, 16);
      xx += __shfl_xor(xx, 32);
      xr = __builtin_sqrtf(xx);
      wave_lds_sync();
    }

    float rbd = FLT_MAX;
    uint32_t rbk = kNone;
    float thi = __builtin_inff();  // min of s_j + M over the centroids screened so far
    for (int c = 0; c < nchunks; ++c) {
      if (nchunks > 1) stage(c * p.KC);
      if (!live) continue;
      const int kc = min(p.KC, K - c * p.KC);
      const float xc = xr + cmx[0];
      const float M = tau2 * (xc * xc);
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
        // Screen.  s_k = |c_k|^2 - 2 x.c_k.  The reference's float distance
        // and |x|^2 + s_k are both within (2D+4) 2^-24 S_k of the exact
        // distance, S_k = (|x| + |c_k|)^2 + |s_k|, so only centroids with
        // s_k - m_k <= s_j + m_j for every j, m_k = tau S_k, can be the
        // reference's argmin.  |s_k| <= (|x| + |c_k|)^2 gives
        // m_k <= M = tau2 (|x| + max |c|)^2, hence the test
        // s_k <= min_j (s_j + M) + M, with j over every centroid screened
        // so far (earlier passes and chunks included).  A NaN s_k, or a NaN
        // or infinite M, passes.
        f4 sv[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const f4 nrm = *reinterpret_cast<const f4 *>(cn + pb + 16 * t + 4 * q);
#pragma unroll
          for (int i = 0; i < 4; ++i) sv[t][i] = __builtin_fmaf(-2.f, acc[t][i], nrm[i]);
        }
        // the lane's entries 4t + i, centroid pb + 16t + 4q + i, that lie in
        // the chunk (every one but in a chunk's last, partial pass)
        uint32_t valid = NT == 8 ? 0xffffffffu : (1u << (4 * NT)) - 1u;
        if (pb + 16 * NT > kc) {  // wave-uniform
          valid = 0;
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int n = min(max(kc - pb - 16 * t - 4 * q, 0), 4);
            valid |= ((1u << n) - 1u) << (4 * t);
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (i >= n) sv[t][i] = __builtin_inff();  // out of the minimum
          }
        }
        float smin = __builtin_inff();
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) smin = sv[t][i] < smin ? sv[t][i] : smin;
        {
          const float o16 = __shfl_xor(smin, 16);
          smin = o16 < smin ? o16 : smin;
          const float o32 = __shfl_xor(smin, 32);
          smin = o32 < smin ? o32 : smin;
        }
        const float th = smin + M;
        thi = th < thi ? th : thi;
        const float lim = thi + M;
        uint32_t cand = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int i = 0; i < 4; ++i) cand |= sv[t][i] > lim ? 0u : 1u << (4 * t + i);
        cand &= valid;
        // exact refinement, ascending k within the lane
        float bd = FLT_MAX;
        uint32_t bk = kNone;
        while (cand) {
          const int b = __builtin_ctz(cand);
          cand &= cand - 1;
          const int kl = pb + 16 * (b >> 2) + 4 * q + (b & 3);
          const float dist = exact_dist<FUSED, NI>(xs + r * DP, cs + kl * DP, nd4);
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

// ---------------------------------------------------------------------------
// Direct form, for narrow rows: each lane takes one row (64 rows per wave)
// and runs the reference's loop itself -- per centroid ascending, per dim
// ascending, one subtraction and one fused (or separately rounded)
// multiply-add -- with the centroids staged in LDS (row stride 4 ceil(D/4),
// zero padded) and read by all lanes at once (broadcast 16-byte reads).
// The wave's 64 x D block of rows is contiguous: it is loaded with
// coalesced 16-byte loads (the next step's while this step computes),
// split into rows in the wave's LDS area (odd row stride: conflict-free)
// and read back one row per lane.  When the centroids exceed one LDS chunk
// the workgroup's waves step through the chunks in lockstep.
// ---------------------------------------------------------------------------
template <int D, bool FUSED>
__global__ __launch_bounds__(64 * kKmDirectWaves) void k_kmeans_direct(KmParams p) {
  constexpr int DS = (D + 3) / 4 * 4;   // staged floats per centroid
  constexpr int XS = D | 1;             // LDS row stride of the row block
  constexpr int NP = (16 * D + 63) / 64;  // 16-byte pieces per lane per block
  extern __shared__ f4 kmd_lds4[];
  float *cs = reinterpret_cast<float *>(kmd_lds4);  // KC x DS
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float *xs = cs + p.KC * DS + w * 64 * XS;  // this wave's 64 rows
  const int K = p.K;
  const int nchunks = (K + p.KC - 1) / p.KC;
  auto stage = [&](int c0) {
    __syncthreads();
    for (int idx = threadIdx.x; idx < p.KC * DS; idx += blockDim.x) {
      const int k = idx / DS, d = idx - k * DS;
      cs[idx] = (c0 + k < K && d < D) ? p.ctrds[(int64_t)(c0 + k) * D + d] : 0.f;
    }
    __syncthreads();
  };
  if (nchunks == 1) stage(0);
  auto row0_of = [&](int64_t st) {
    return ((st * gridDim.x + blockIdx.x) * kKmDirectWaves + w) * 64;
  };
  // piece f = lane + 64 i of the block: floats 4f .. 4f+3 (past the end of
  // the array: zero; a piece that crosses it element by element)
  auto load_block = [&](int64_t st, f4 (&dst)[NP]) {
    const int64_t row0 = row0_of(st);
    const int64_t left = row0 < p.nrows ? (p.nrows - row0) * D * 4 : 0;
    const __amdgpu_buffer_rsrc_t rs =
        make_rsrc(p.rows + (row0 < p.nrows ? row0 : 0) * D, left < kRsrcMax ? (int32_t)left : kRsrcMax);
#pragma unroll
    for (int i = 0; i < NP; ++i) {
      const int64_t o = 16 * (int64_t)(lane + 64 * i);
      if (lane + 64 * i >= 16 * D || o >= left) {
        dst[i] = f4{0.f, 0.f, 0.f, 0.f};
      } else if (o + 16 <= left) {
        dst[i] = buf_load4<kBufNT>(rs, (uint32_t)o, 0);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) dst[i][e] = buf_load1(rs, o + 4 * e < left ? (uint32_t)(o + 4 * e) : kOOB, 0);
      }
    }
  };
  f4 pnext[NP];
  load_block(0, pnext);
  for (int64_t st = 0; st < p.steps; ++st) {
    const int64_t row0 = row0_of(st);
    const bool live = row0 < p.nrows;  // wave-uniform
    if (nchunks == 1 && !live) break;
    float x[DS];
    if (live) {
      // split the block into rows: element 4f + e is row (4f+e)/D, dim (4f+e)%D
#pragma unroll
      for (int i = 0; i < NP; ++i) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int idx = 4 * (lane + 64 * i) + e;
          if (idx < 64 * D) xs[(idx / D) * XS + idx % D] = pnext[i][e];
        }
      }
      wave_lds_sync();
#pragma unroll
      for (int d = 0; d < DS; ++d) x[d] = d < D ? xs[lane * XS + d] : 0.f;
      wave_lds_sync();  // the next step's split may overwrite the area
    }
    if (st + 1 < p.steps) load_block(st + 1, pnext);
    const bool rin = row0 + lane < p.nrows;
    float best = FLT_MAX;
    uint32_t bk = kNone;
    for (int c = 0; c < nchunks; ++c) {
      if (nchunks > 1) stage(c * p.KC);
      if (!live) continue;
      const int kc = min(p.KC, K - c * p.KC);
      for (int k = 0; k < kc; ++k) {
        const float *cr = cs + k * DS;
        float dist = 0.f;
#pragma unroll
        for (int j = 0; j < DS / 4; ++j) {
          const f4 cv = *reinterpret_cast<const f4 *>(cr + 4 * j);
#pragma unroll
          for (int e = 0; e < 4; ++e) dist = exact_step<FUSED>(dist, x[4 * j + e], cv[e]);
        }
        if (dist < best) {
          best = dist;
          bk = (uint32_t)(c * p.KC + k);
        }
      }
    }
    if (live && rin) p.assign[row0 + lane] = bk == kNone ? (uint32_t)K : bk;
  }
}

template <int D>
KmKernel kmd_pick(bool fused) {
  return fused ? k_kmeans_direct<D, true> : k_kmeans_direct<D, false>;
}

template <int... Ds>
KmKernel kmd_find(int d, bool fused, std::integer_sequence<int, Ds...>) {
  KmKernel out = nullptr;
  ((d == Ds + 1 ? (out = kmd_pick<Ds + 1>(fused), 0) : 0), ...);
  return out;
}

KmKernel kmd_kernel(int d, bool fused) {
  return kmd_find(d, fused, std::make_integer_sequence<int, kKmDirectMaxDims>{});
}

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

KmeansPlan plan_kmeans(int64_t nrows, int D, int K, int mode) {
  KmeansPlan pl;
  // measured on MI355X (2^24 rows, dims 4..32 x clusters 4..1024): the
  // direct kernel wins up to 8 dims at any cluster count and below
  // clusters x dims = 1024 up to 32 dims (by up to 9x at few clusters);
  // the screened kernel wins above
  pl.direct = (mode == 2 && D <= kKmDirectMaxDims) ||
              (mode == 0 && D <= kKmDirectMaxDims && (D <= 8 || (int64_t)K * D < 1024));
  if (pl.direct) {
    // one row per lane; centroids in chunks of at most kKmDirectLdsBytes,
    // plus each wave's 64-row area
    pl.ni = D;
    const int64_t ds = (D + 3) / 4 * 4;
    const int64_t rows_f = (int64_t)kKmDirectWaves * 64 * (D | 1);
    int64_t kc = (int64_t)(kKmDirectLdsBytes / 4) / ds;
    if (kc > K) kc = K;
    pl.KC = (int)kc;
    pl.lds = (size_t)(kc * ds + rows_f) * 4;
    pl.tiles = (nrows + 63) / 64;  // 64-row wave blocks
    return pl;
  }
  pl.ni = (D + 15) / 16;
  pl.nt = K <= 16 ? 1 : K <= 32 ? 2 : K <= 64 ? 4 : 8;
  const int dp = 16 * pl.ni + 4;
  const int pass = 16 * pl.nt;
  // LDS: centroid chunk (KC x dp) + squared norms (KC) + the chunk's
  // largest norm (4 floats, keeping 16-byte alignment) + 16 rows per wave,
  // at most kKmLdsBytes.
  const int64_t rows_f = (int64_t)kKmWaves * 16 * dp;
  const int64_t room = kKmLdsBytes / 4 - rows_f - 4;
  int64_t kc = room / (dp + 1) / pass * pass;
  const int64_t kpad = ((int64_t)K + pass - 1) / pass * pass;
  if (kc > kpad) kc = kpad;
  pl.KC = (int)kc;
  pl.lds = (size_t)(kc * (dp + 1) + 4 + rows_f) * 4;
  pl.tiles = (nrows + 15) / 16;
  return pl;
}

hipError_t launch_kmeans_assign(const KmeansPlan &pl, uint32_t *assign, const float *rows,
                                int64_t nrows, int D, const float *ctrds, int K, bool fused,
                                hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  if (pl.direct ? pl.KC < 1 : pl.KC < 16 * pl.nt) return hipErrorInvalidValue;
  KmKernel fn = pl.direct ? kmd_kernel(pl.ni, fused) : km_kernel(pl.ni, pl.nt, fused);
  if (!fn) return hipErrorInvalidValue;
  const int waves = pl.direct ? kKmDirectWaves : kKmWaves;
  hipError_t e;
  if (pl.lds > 65536) {
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds);
    if (e != hipSuccess) return e;
  }
  int per_cu = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(fn),
                                                   64 * waves, pl.lds);
  if (e != hipSuccess) return e;
  if (per_cu < 1) per_cu = 1;
  const int64_t wgs_needed = (pl.tiles + waves - 1) / waves;
  int64_t grid = (int64_t)per_cu * cu_count();
  if (grid > wgs_needed) grid = wgs_needed;
  KmParams p;
  p.rows = rows;
  p.ctrds = ctrds;
  p.assign = assign;
  p.nrows = nrows;
  p.steps = (pl.tiles + grid * waves - 1) / (grid * waves);
  p.D = D;
  p.K = K;
  p.KC = pl.KC;
  hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(64 * waves), pl.lds, st, p);
  return hipGetLastError();
}

}  // namespace hml
