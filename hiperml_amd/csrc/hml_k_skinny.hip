// hml_k_skinny.hip -- skinny SGEMM on gfx950 FP32 matrix cores.
//
// Replaces the reference's register-blocked CUDA tile kernels
// hmlSgemmVarKNN / hmlSgemmConstKNN (include/hml_sgemm_kernel_template.h:
// 19-191, 229-333).  Same operation C = alpha*A*B + beta*C; different
// machine model:
//
//  * The streamed operand A (the big one) is read exactly once from HBM with
//    128-bit loads straight into registers; no LDS staging, no barriers.
//    Every wave is an independent work item (a block of rows x a K range).
//  * The small operand B is packed once per call into the fragment order of
//    v_mfma_f32_16x16x4f32, so each wave fetches its B fragments with
//    coalesced 128-bit loads that hit in L1/L2 (B is re-read by every row
//    block; it is small by construction).
//  * FP32 products run on the matrix cores (2x the non-packed VALU rate).
//    v_mfma_f32_16x16x4f32: lane l supplies A[m=l%16][k=l/16] and
//    B[k=l/16][n=l%16]; the 16x16 result sits in 4 VGPRs, lane l holding
//    D[4*(l/16)+i][l%16].  The reduction index k is a sum, so any
//    permutation of K that A and B share is legal; each form picks the
//    permutation that makes its A loads contiguous.
//  * K is split across waves when rows alone cannot fill 256 CUs; the
//    partial tiles are summed in a fixed order by a second kernel
//    (deterministic, no atomics).
#include "hml_internal.h"

namespace hml {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Loads 4 consecutive floats p[0..3] of which the first `nvalid` are in
// range (0..4); the rest read as zero.  `vec` says p is 16-byte aligned.
template <bool VEC>
__device__ __forceinline__ f4 load4(const float *p, int64_t nvalid) {
  f4 v = {0.f, 0.f, 0.f, 0.f};
  if (nvalid >= 4) {
    if (VEC) {
      v = *reinterpret_cast<const f4 *>(p);
    } else {
      v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
    }
  } else if (nvalid > 0) {
    v[0] = p[0];
    if (nvalid > 1) v[1] = p[1];
    if (nvalid > 2) v[2] = p[2];
  }
  return v;
}

// Unguarded load of 4 consecutive floats that are read exactly once
// (non-temporal, so the streamed operand does not evict B from the caches).
template <bool VEC>
__device__ __forceinline__ f4 stream4(const float *p) {
  if (VEC) return __builtin_nontemporal_load(reinterpret_cast<const f4 *>(p));
  f4 v;
  v[0] = __builtin_nontemporal_load(p);
  v[1] = __builtin_nontemporal_load(p + 1);
  v[2] = __builtin_nontemporal_load(p + 2);
  v[3] = __builtin_nontemporal_load(p + 3);
  return v;
}

__device__ __forceinline__ void store_out(const SkinnyParams &p, int64_t row,
                                          int64_t col, int ks, float acc) {
  if (row < p.M && col < p.N) {
    if (p.ksplit == 1) {
      float *cp = p.C + row * p.sCm + col * p.sCn;
      float v = p.alpha * acc;
      if (p.read_c) v = p.alpha * acc + p.beta * (*cp);
      *cp = v;
    } else {
      p.P[((int64_t)ks * p.M + row) * p.N + col] = acc;
    }
  }
}

// Both kernels stream K in steps.  Steps that lie wholly inside the wave's
// K range go through a guard-free pipeline that keeps P steps of loads in
// flight in P named register buffers (the loop is unrolled by P, so no
// buffer is ever copied and the compiler waits only for the oldest step);
// a partial last step (K not a multiple of the step) is loaded with bounds
// checks afterwards.

// ---------------------------------------------------------------------------
// DOT form: A(m,k) = A[m*sAm + k].  Per k-step of 32, lane (g = l/16,
// r = l%16) of row tile t loads A(row_t + r, k0 + 16v + 4g .. +3) for
// v = 0, 1: each 128-bit load instruction reads 64 contiguous bytes of each
// of its 16 rows.  MFMA number 4v+j consumes element j of load v, so the
// shared K permutation is k = k0 + 16v + 4g + j.  The packed B fragment of
// (step s, column tile c, half v) holds, for lane (g, r),
// B(32s + 16v + 4g + j, 16c + r) for j = 0..3.
// ---------------------------------------------------------------------------
// One wave computes rows [rb*16*RT, (rb+1)*16*RT) x column tiles
// [ct0, ct0+CT) over the K range of split ks.
template <int RT, int CT, bool VEC>
__device__ __forceinline__ void dot_tile(const SkinnyParams &p, int64_t rb,
                                         int ks, int ct0) {
  constexpr int P = (RT + CT <= 3) ? 4 : 2;
  const int lane = threadIdx.x & 63;
  const int64_t rows_per_wave = (int64_t)kDotRowsPerTile * RT;
  const int r = lane & 15, g = lane >> 4;

  const float *arow[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    int64_t row = rb * rows_per_wave + t * 16 + r;
    if (row > p.M - 1) row = p.M - 1;  // clamp: loads stay in range, stores are masked
    arow[t] = p.A + row * p.sAm;
  }
  const int64_t kb = (int64_t)ks * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;
  const int64_t nfull = ke > kb ? (ke - kb) / kDotKStep : 0;

  const f4 *Bp = reinterpret_cast<const f4 *>(p.Bp);
  const int64_t bstep = (int64_t)p.nct_pad * 128;  // f4 per k-step
  const f4 *bp = Bp + (kb / kDotKStep) * bstep + (int64_t)ct0 * 128 + lane;
  const float *ap[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) ap[t] = arow[t] + kb + 4 * g;

  f4 acc[RT][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = f4{0.f, 0.f, 0.f, 0.f};

  f4 a[P][RT][2], b[P][CT][2];
  auto load = [&](int j, int64_t s) {
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      a[j][t][0] = stream4<VEC>(ap[t] + s * kDotKStep);
      a[j][t][1] = stream4<VEC>(ap[t] + s * kDotKStep + 16);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      b[j][c][0] = bp[s * bstep + c * 128];
      b[j][c][1] = bp[s * bstep + c * 128 + 64];
    }
  };
  auto comp = [&](int j) {
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[t][c] = mfma4(a[j][t][v][jj], b[j][c][v][jj], acc[t][c]);
  };

  if (nfull >= P) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      load(j, j);
      __builtin_amdgcn_sched_barrier(0);  // issue the stages in order
    }
    int64_t s = 0;
    for (; s + 2 * P <= nfull; s += P) {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        comp(j);
        load(j, s + P + j);
        __builtin_amdgcn_sched_barrier(0);  // keep each stage's loads right after its compute
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      comp(j);
      if (s + P + j < nfull) load(j, s + P + j);
    }
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (s + P + j < nfull) comp(j);
  } else {
    for (int64_t s = 0; s < nfull; ++s) {
      load(0, s);
      comp(0);
    }
  }
  if (kb + nfull * kDotKStep < ke) {
    // partial last step: guarded A, B is zero padded past K
    const int64_t k0 = kb + nfull * kDotKStep, kk = k0 + 4 * g;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      a[0][t][0] = load4<VEC>(arow[t] + kk, ke - kk);
      a[0][t][1] = load4<VEC>(arow[t] + kk + 16, ke - kk - 16);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      b[0][c][0] = bp[nfull * bstep + c * 128];
      b[0][c][1] = bp[nfull * bstep + c * 128 + 64];
    }
    comp(0);
  }

#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = rb * rows_per_wave + t * 16 + 4 * g + i;
        const int64_t col = (int64_t)(ct0 + c) * 16 + r;
        store_out(p, row, col, ks, acc[t][c][i]);
      }
}

template <int RT, int CT, bool VEC>
__global__ __launch_bounds__(256) void k_skinny_dot(SkinnyParams p) {
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)kDotRowsPerTile * RT;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.ksplit;
  const int ks = (int)(task - rb * p.ksplit);
  if (rb >= nrb) return;  // whole wave leaves; the kernel has no barriers
  dot_tile<RT, CT, VEC>(p, rb, ks, blockIdx.y * CT);
}

// ---------------------------------------------------------------------------
// Persistent grouped SGEMM: hundreds of independent skinny problems in one
// launch.  The host builds one descriptor per problem and a list of
// (problem, row block) tiles sorted by decreasing cost; wave w of the
// persistent grid takes tiles w, w + W, w + 2W, ... (W = waves launched), a
// static longest-first schedule that needs no atomics or counters, so the
// launch can be captured in a graph and replayed.  Each tile runs the same
// pipelined DOT body as the single-problem kernel (16 rows x every column
// of its problem, whole K, no split), dispatching on the problem's column
// tile count and alignment.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grouped_dot(const GroupProblem *probs,
                                                     const GroupTile *tiles,
                                                     int64_t ntiles) {
  const int64_t W = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  for (int64_t t = w; t < ntiles; t += W) {
    const GroupTile tile = tiles[t];
    const GroupProblem &gp = probs[tile.prob];
    SkinnyParams p{};
    p.A = gp.A; p.sAm = gp.lda; p.sAk = 1;
    p.Bp = gp.Bp;
    p.C = gp.C; p.sCm = gp.ldc; p.sCn = 1;
    p.P = nullptr;
    p.M = gp.M; p.N = gp.N; p.K = gp.K;
    p.kchunk = gp.K > 0 ? gp.K : 1; p.ksplit = 1;
    p.nct_pad = gp.nct_pad;
    p.alpha = gp.alpha; p.beta = gp.beta; p.read_c = gp.read_c;
    const int kind = gp.ct * 2 + gp.vec;
    for (int ct0 = 0; ct0 < gp.nct_pad; ct0 += gp.ct) {
      switch (kind) {
        case 2: dot_tile<1, 1, false>(p, tile.rb, 0, ct0); break;
        case 3: dot_tile<1, 1, true>(p, tile.rb, 0, ct0); break;
        case 4: dot_tile<1, 2, false>(p, tile.rb, 0, ct0); break;
        case 5: dot_tile<1, 2, true>(p, tile.rb, 0, ct0); break;
        case 8: dot_tile<1, 4, false>(p, tile.rb, 0, ct0); break;
        case 9: dot_tile<1, 4, true>(p, tile.rb, 0, ct0); break;
        default: break;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// AXPY form: A(m,k) = A[m + k*sAk].  A row tile is 64 rows: lane (g, r)
// loads A(m0 + 4r .. m0 + 4r + 3, k) for k = k0 + 4u + g (u = 0..3 per
// k-step of 16).  MFMA j (j = 0..3) takes element j, so its 16 "rows" are
// the interleaved rows m0 + 4r + j; its B operand B(k0 + 4u + g, 16c + r)
// is shared by the 4 MFMAs.  Packed B fragment (step s, tile c) holds, for
// lane (g, r), B(16s + 4u + g, 16c + r) for u = 0..3.
// VEC requires M % 4 == 0, so a lane's 4 rows are all in range or all out;
// out-of-range lanes load row 0 and multiply zeros.  The scalar variant
// clamps and zeroes each element on its own.
// ---------------------------------------------------------------------------
template <int RT, int CT, bool VEC>
__global__ __launch_bounds__(256) void k_skinny_axpy(SkinnyParams p) {
  constexpr int P = (RT == 1 && CT <= 2) ? 3 : 2;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)kAxpyRowsPerTile * RT;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.ksplit;
  const int ks = (int)(task - rb * p.ksplit);
  if (rb >= nrb) return;
  const int ct0 = blockIdx.y * CT;
  const int r = lane & 15, g = lane >> 4;

  // per lane: 4 rows m4[t] .. m4[t] + 3 of row tile t
  int64_t m4[RT];
  bool rin[RT][4];
  const float *arow[RT][4];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    m4[t] = rb * rows_per_wave + t * 64 + 4 * r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      rin[t][e] = m4[t] + e < p.M;
      arow[t][e] = p.A + (rin[t][e] ? m4[t] + e : (VEC ? 0 : p.M - 1));
    }
  }
  const int64_t kb = (int64_t)ks * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;
  const int64_t nfull = ke > kb ? (ke - kb) / kAxpyKStep : 0;

  const f4 *Bp = reinterpret_cast<const f4 *>(p.Bp);
  const int64_t bstep = (int64_t)p.nct_pad * 64;
  const f4 *bp = Bp + (kb / kAxpyKStep) * bstep + (int64_t)ct0 * 64 + lane;

  f4 acc[RT][4][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][j][c] = f4{0.f, 0.f, 0.f, 0.f};

  // A values of row tile t at column k (zero where the row is outside M)
  auto ldA = [&](int t, int64_t k) -> f4 {
    f4 v;
    if (VEC) {
      v = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(arow[t][0] + k * p.sAk));
      if (!rin[t][0]) v = f4{0.f, 0.f, 0.f, 0.f};
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float x = __builtin_nontemporal_load(arow[t][e] + k * p.sAk);
        v[e] = rin[t][e] ? x : 0.f;
      }
    }
    return v;
  };

  f4 a[P][RT][4], b[P][CT];
  auto load = [&](int j, int64_t s) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = kb + s * kAxpyKStep + 4 * u + g;
#pragma unroll
      for (int t = 0; t < RT; ++t) a[j][t][u] = ldA(t, k);
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) b[j][c] = bp[s * bstep + c * 64];
  };
  auto comp = [&](int j) {
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[t][jj][c] = mfma4(a[j][t][u][jj], b[j][c][u], acc[t][jj][c]);
  };

  if (nfull >= P) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
      load(j, j);
      __builtin_amdgcn_sched_barrier(0);  // issue the stages in order
    }
    int64_t s = 0;
    for (; s + 2 * P <= nfull; s += P) {
#pragma unroll
      for (int j = 0; j < P; ++j) {
        comp(j);
        load(j, s + P + j);
        __builtin_amdgcn_sched_barrier(0);  // keep each stage's loads right after its compute
      }
    }
#pragma unroll
    for (int j = 0; j < P; ++j) {
      comp(j);
      if (s + P + j < nfull) load(j, s + P + j);
    }
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (s + P + j < nfull) comp(j);
  } else {
    for (int64_t s = 0; s < nfull; ++s) {
      load(0, s);
      comp(0);
    }
  }
  if (kb + nfull * kAxpyKStep < ke) {
    // partial last step: k beyond ke reads as zero
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = kb + nfull * kAxpyKStep + 4 * u + g;
#pragma unroll
      for (int t = 0; t < RT; ++t)
        a[0][t][u] = k < ke ? ldA(t, k) : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) b[0][c] = bp[nfull * bstep + c * 64];
    comp(0);
  }

#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int64_t row = rb * rows_per_wave + t * 64 + 4 * (4 * g + i) + j;
          const int64_t col = (int64_t)(ct0 + c) * 16 + r;
          store_out(p, row, col, ks, acc[t][j][c][i]);
        }
}

// ---------------------------------------------------------------------------
// B packing.  One thread writes one 128-bit fragment element.
// ---------------------------------------------------------------------------
__device__ __forceinline__ f4 pack_elem(const float *__restrict__ B,
                                        int64_t sBk, int64_t sBn, int64_t K,
                                        int64_t N, int form, int nct_pad,
                                        int64_t idx) {
  const int lane = (int)(idx & 63);
  const int r = lane & 15, g = lane >> 4;
  f4 v = {0.f, 0.f, 0.f, 0.f};
  int64_t n, kbase, kstride;
  if (form == kFormDot) {
    const int64_t q = idx >> 6;         // ((s*nct_pad + c)*2 + half)
    const int half = (int)(q & 1);
    const int64_t sc = q >> 1;
    const int64_t c = sc % nct_pad, s = sc / nct_pad;
    n = c * 16 + r;
    kbase = s * kDotKStep + 16 * half + 4 * g;
    kstride = 1;
  } else {
    const int64_t sc = idx >> 6;        // s*nct_pad + c
    const int64_t c = sc % nct_pad, s = sc / nct_pad;
    n = c * 16 + r;
    kbase = s * kAxpyKStep + g;
    kstride = 4;
  }
  if (n < N) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t k = kbase + j * kstride;
      if (k < K) v[j] = B[k * sBk + n * sBn];
    }
  }
  return v;
}

__global__ void k_pack_b(const float *__restrict__ B, int64_t sBk, int64_t sBn,
                         int64_t K, int64_t N, int form, int64_t total,
                         int nct_pad, f4 *__restrict__ Bp) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  Bp[idx] = pack_elem(B, sBk, sBn, K, N, form, nct_pad, idx);
}

// blockIdx.y = problem; blocks past a problem's packed size exit.
__global__ void k_pack_group(const GroupPack *__restrict__ packs) {
  const GroupPack pk = packs[blockIdx.y];
  const int64_t nsteps = (pk.K + kDotKStep - 1) / kDotKStep;
  const int64_t total = nsteps * pk.nct_pad * 128;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  reinterpret_cast<f4 *>(pk.Bp)[idx] =
      pack_elem(pk.B, pk.sBk, pk.sBn, pk.K, pk.N, kFormDot, pk.nct_pad, idx);
}

// ---------------------------------------------------------------------------
// Split reduction.  "wide": one thread per output, serial over R (small R).
// "tall": 16 waves share 64 outputs and split R, then combine in LDS in a
// fixed order (large R, few outputs).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void epilogue_store(float *C, int64_t idx,
                                               int64_t ninner, int64_t sCm,
                                               int64_t sCn, float alpha,
                                               float beta, int read_c,
                                               float s) {
  const int64_t m = idx / ninner, n = idx - m * ninner;
  float *cp = C + m * sCm + n * sCn;
  float v = alpha * s;
  if (read_c) v = alpha * s + beta * (*cp);
  *cp = v;
}

__global__ void k_reduce_wide(const float *__restrict__ P, int R,
                              int64_t total, int64_t ninner, float *C,
                              int64_t sCm, int64_t sCn, float alpha,
                              float beta, int read_c) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  float s = 0.f;
  for (int r = 0; r < R; ++r) s += P[(int64_t)r * total + idx];
  epilogue_store(C, idx, ninner, sCm, sCn, alpha, beta, read_c, s);
}

__global__ __launch_bounds__(1024) void k_reduce_tall(
    const float *__restrict__ P, int R, int64_t total, int64_t ninner,
    float *C, int64_t sCm, int64_t sCn, float alpha, float beta, int read_c) {
  __shared__ float part[16][64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t idx = (int64_t)blockIdx.x * 64 + l;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (idx < total) {
    int r = w;
    for (; r + 48 < R; r += 64) {
      s0 += P[(int64_t)r * total + idx];
      s1 += P[(int64_t)(r + 16) * total + idx];
      s2 += P[(int64_t)(r + 32) * total + idx];
      s3 += P[(int64_t)(r + 48) * total + idx];
    }
    for (; r < R; r += 16) s0 += P[(int64_t)r * total + idx];
  }
  part[w][l] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (w == 0 && idx < total) {
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) s += part[i][l];
    epilogue_store(C, idx, ninner, sCm, sCn, alpha, beta, read_c, s);
  }
}

__global__ void k_scale(float *C, int64_t M, int64_t N, int64_t sCm,
                        int64_t sCn, float alpha, float beta, int read_c) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= M * N) return;
  const int64_t m = idx / N, n = idx - m * N;
  float *cp = C + m * sCm + n * sCn;
  const float s = alpha * 0.0f;
  *cp = read_c ? s + beta * (*cp) : s;
}

__global__ void k_fill(float *dst, size_t n, float value) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = value;
}

template <int RT, int CT, bool VEC>
hipError_t run_dot(int wpb, const SkinnyParams &p, hipStream_t st) {
  const int64_t rows = (int64_t)kDotRowsPerTile * RT;
  const int64_t tasks = ((p.M + rows - 1) / rows) * p.ksplit;
  const int64_t nblk = (tasks + wpb - 1) / wpb;
  const int64_t ncb = p.nct_pad / CT;
  if (nblk > 0x7fffffff || ncb > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_skinny_dot<RT, CT, VEC>), dim3((unsigned)nblk, (unsigned)ncb),
                     dim3(64 * wpb), 0, st, p);
  return hipGetLastError();
}

template <int RT, int CT, bool VEC>
hipError_t run_axpy(int wpb, const SkinnyParams &p, hipStream_t st) {
  const int64_t rows = (int64_t)kAxpyRowsPerTile * RT;
  const int64_t tasks = ((p.M + rows - 1) / rows) * p.ksplit;
  const int64_t nblk = (tasks + wpb - 1) / wpb;
  const int64_t ncb = p.nct_pad / CT;
  if (nblk > 0x7fffffff || ncb > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_skinny_axpy<RT, CT, VEC>), dim3((unsigned)nblk, (unsigned)ncb),
                     dim3(64 * wpb), 0, st, p);
  return hipGetLastError();
}

typedef hipError_t (*RunFn)(int, const SkinnyParams &, hipStream_t);

template <int RT, int CT>
constexpr RunFn dot_fn(bool vec) { return vec ? run_dot<RT, CT, true> : run_dot<RT, CT, false>; }
template <int RT, int CT>
constexpr RunFn axpy_fn(bool vec) { return vec ? run_axpy<RT, CT, true> : run_axpy<RT, CT, false>; }

RunFn find_skinny(int form, int rt, int ct, bool vec) {
#define HML_CT_CASES(FN, RT)                    \
  switch (ct) {                                 \
    case 1: return FN<RT, 1>(vec);              \
    case 2: return FN<RT, 2>(vec);              \
    case 4: return FN<RT, 4>(vec);              \
    case 8: return FN<RT, 8>(vec);              \
    default: return nullptr;                    \
  }
  if (form == kFormDot) {
    switch (rt) {
      case 1: HML_CT_CASES(dot_fn, 1)
      case 2: HML_CT_CASES(dot_fn, 2)
      case 4: HML_CT_CASES(dot_fn, 4)
      default: return nullptr;
    }
  } else if (form == kFormAxpy) {
    switch (rt) {
      case 1: HML_CT_CASES(axpy_fn, 1)
      case 2:
        if (ct == 8) return nullptr;  // 2 x 64 rows x 128 columns spills
        HML_CT_CASES(axpy_fn, 2)
      default: return nullptr;
    }
  }
#undef HML_CT_CASES
  return nullptr;
}

}  // namespace

bool skinny_has_variant(int form, int rt, int ct) {
  return find_skinny(form, rt, ct, true) != nullptr;
}

hipError_t launch_skinny(int form, int rt, int ct, bool vec, int wpb,
                         const SkinnyParams &p, hipStream_t st) {
  RunFn fn = find_skinny(form, rt, ct, vec);
  if (!fn || (wpb != 1 && wpb != 2 && wpb != 4)) return hipErrorInvalidValue;
  if (p.nct_pad % ct != 0 || p.ksplit < 1) return hipErrorInvalidValue;
  return fn(wpb, p, st);
}

size_t packed_b_floats(int form, int64_t K, int nct_pad) {
  const int64_t step = form == kFormDot ? kDotKStep : kAxpyKStep;
  const int64_t nsteps = (K + step - 1) / step;
  // DOT: 2 fragments of 4 floats per lane per (step, tile); AXPY: 1.
  return (size_t)nsteps * nct_pad * 64 * 4 * (form == kFormDot ? 2 : 1);
}

hipError_t launch_pack_b(const float *B, int64_t sBk, int64_t sBn, int64_t K,
                         int64_t N, int form, int64_t nsteps, int nct_pad,
                         float *Bp, hipStream_t st) {
  const int64_t total = nsteps * nct_pad * 64 * (form == kFormDot ? 2 : 1);
  if (total <= 0) return hipSuccess;
  const int64_t nblk = (total + 255) / 256;
  hipLaunchKernelGGL(k_pack_b, dim3((unsigned)nblk), dim3(256), 0, st, B, sBk,
                     sBn, K, N, form, total, nct_pad,
                     reinterpret_cast<f4 *>(Bp));
  return hipGetLastError();
}

hipError_t launch_reduce(const float *P, int R, int64_t total, int64_t ninner,
                         float *C, int64_t sCm, int64_t sCn, float alpha,
                         float beta, int read_c, hipStream_t st) {
  if (total <= 0) return hipSuccess;
  if (R > 32 && total <= (int64_t)64 * 2048) {
    const int64_t nblk = (total + 63) / 64;
    hipLaunchKernelGGL(k_reduce_tall, dim3((unsigned)nblk), dim3(1024), 0, st,
                       P, R, total, ninner, C, sCm, sCn, alpha, beta, read_c);
  } else {
    const int64_t nblk = (total + 255) / 256;
    hipLaunchKernelGGL(k_reduce_wide, dim3((unsigned)nblk), dim3(256), 0, st,
                       P, R, total, ninner, C, sCm, sCn, alpha, beta, read_c);
  }
  return hipGetLastError();
}

hipError_t launch_scale(float *C, int64_t M, int64_t N, int64_t sCm,
                        int64_t sCn, float alpha, float beta, int read_c,
                        hipStream_t st) {
  const int64_t total = M * N;
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, st, C, M, N, sCm, sCn, alpha, beta, read_c);
  return hipGetLastError();
}

hipError_t launch_fill(float *dst, size_t n, float value, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t nblk = (n + 255) / 256;
  if (nblk > 65536) nblk = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)nblk), dim3(256), 0, st, dst, n,
                     value);
  return hipGetLastError();
}

}  // namespace hml

namespace hml {

hipError_t launch_grouped(const GroupProblem *probs, const GroupTile *tiles,
                          int64_t ntiles, int nblocks, hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  if (nblocks < 1) nblocks = 1;
  hipLaunchKernelGGL(k_grouped_dot, dim3((unsigned)nblocks), dim3(256), 0, st,
                     probs, tiles, ntiles);
  return hipGetLastError();
}

hipError_t launch_pack_group(const GroupPack *packs, int32_t count,
                             int64_t max_elems, hipStream_t st) {
  if (count <= 0 || max_elems <= 0) return hipSuccess;
  const int64_t nblk = (max_elems + 255) / 256;
  if (nblk > 0x7fffffff || count > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_pack_group, dim3((unsigned)nblk, (unsigned)count),
                     dim3(256), 0, st, packs);
  return hipGetLastError();
}

}  // namespace hml
