// hml_k_gemv.hip -- SGEMV on gfx950 (wave64, 128-bit loads, DPP reductions).
//
// Replaces the reference's three CUDA GEMV families BasicN / VarN / ConstNN
// (include/hml_sgemv_kernel_template.h:19-106, 143-290, 328-411), which stage
// A through shared memory and reduce with warp-synchronous volatile stores.
// Here two forms cover every layout:
//
//  DOT  : y[m] = sum_n A[m*lda + n] x[n]  (each output a contiguous row).
//         A wave splits into 64/LPR lane groups; a group streams one row
//         with 128-bit loads, LPR lanes side by side, and reduces with DPP
//         row operations (quad_perm, row_half_mirror, row_mirror) plus two
//         cross-row shuffles.  Each lane group handles RB rows so one x load
//         serves RB rows.  Optional split of n across waves: the waves of a
//         workgroup sum their chunks in LDS, partials only for splits wider
//         than a workgroup.
//  AXPY : y[n] = sum_m A[m*lda + n] x[m]  (each output a column sum).
//         LPR lanes x 4 floats span a column chunk; the 64/LPR lane groups
//         and the waves of a block take interleaved rows; lane groups fold
//         with shuffles, waves through LDS, and row splits through a
//         fixed-order partial reduction (no atomics, deterministic).
//
// Neither form relies on out-of-range reads returning zero (the reference
// relies on texture clamping, include/hml_sgemv_kernel_template.h:193-211):
// every access is bounds-checked.
#include "hml_internal.h"
#include "hml_kdev.h"

namespace hml {
namespace {

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over aligned groups of LPR consecutive lanes; every lane of a group
// ends with the group total.
template <int LPR>
__device__ __forceinline__ float group_sum(float v) {
  if (LPR >= 2) v += dpp<0xB1>(v);    // quad_perm [1,0,3,2]
  if (LPR >= 4) v += dpp<0x4E>(v);    // quad_perm [2,3,0,1]
  if (LPR >= 8) v += dpp<0x141>(v);   // row_half_mirror
  if (LPR >= 16) v += dpp<0x140>(v);  // row_mirror
  if (LPR >= 32) v += __shfl_xor(v, 16);
  if (LPR >= 64) v += __shfl_xor(v, 32);
  return v;
}

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

// End of a DOT wave: its lane groups' sums of RB rows (row0 + iG + grp) go
// to y, or to slice `sp` of the partials.  With p.wsum the waves of the
// workgroup hold consecutive chunks of the same rows: they meet in LDS and
// wave 0 adds them in wave order, writing slice sp / wpb (no partials at all
// when one workgroup covers the whole row).
template <int LPR, int RB>
__device__ __forceinline__ void dot_finish(const GemvParams &p, const float (&acc)[RB],
                                           int64_t row0, int sp) {
  constexpr int G = 64 / LPR;
  __shared__ float red[kGemvMaxWaves * RB * G];
  const int lane = threadIdx.x & 63;
  const int grp = lane / LPR, q = lane % LPR;
  float s[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) s[i] = group_sum<LPR>(acc[i]);
  int slot = sp, nslot = p.split;
  if (p.wsum) {
    const int w = threadIdx.x >> 6, wpb = blockDim.x >> 6;
    if (q == 0) {
#pragma unroll
      for (int i = 0; i < RB; ++i) red[(w * RB + i) * G + grp] = s[i];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      float t = red[i * G + grp];
      for (int ww = 1; ww < wpb; ++ww) t += red[(ww * RB + i) * G + grp];
      s[i] = t;
    }
    slot = sp / wpb;
    nslot = p.split / wpb;
  }
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int64_t row = row0 + (int64_t)i * G + grp;
    if (q == 0 && row < p.M) {
      if (nslot == 1) {
        float *yp = p.y + row;
        float v = p.alpha * s[i];
        if (p.read_c) v = p.alpha * s[i] + p.beta * (*yp);
        *yp = v;
      } else {
        p.P[(int64_t)slot * p.M + row] = s[i];
      }
    }
  }
}

template <int LPR, int RB, bool VEC>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_dot(GemvParams p) {
  constexpr int G = 64 / LPR;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)G * RB;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.split;
  const int sp = (int)(task - rb * p.split);
  if (rb >= nrb) return;  // with p.wsum the whole workgroup shares rb
  const int grp = lane / LPR, q = lane % LPR;

  const float *arow[RB];
  int64_t row[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    row[i] = rb * rows_per_wave + (int64_t)i * G + grp;
    const int64_t rc = row[i] < p.M ? row[i] : p.M - 1;
    arow[i] = p.A + rc * p.lda;
  }
  const int64_t cb = (int64_t)sp * p.chunk;
  int64_t ce = cb + p.chunk;
  if (ce > p.N) ce = p.N;

  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.f;

  constexpr int STEP = 4 * LPR;
#pragma unroll 2
  for (int64_t c = cb + 4 * q; c < ce; c += STEP) {
    const f4 xv = load4<VEC>(p.x + c, ce - c);
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const f4 av = load4<VEC>(arow[i] + c, ce - c);
      acc[i] += av[0] * xv[0];
      acc[i] += av[1] * xv[1];
      acc[i] += av[2] * xv[2];
      acc[i] += av[3] * xv[3];
    }
  }

  dot_finish<LPR, RB>(p, acc, rb * rows_per_wave, sp);
}

template <int LPR, bool VEC>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_axpy(GemvParams p) {
  constexpr int G = 64 / LPR;
  constexpr int CW = 4 * LPR;
  __shared__ f4 red[kGemvMaxWaves][LPR];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int grp = lane / LPR, q = lane % LPR;
  const int64_t c0 = (int64_t)blockIdx.x * CW + 4 * q;
  const int sp = blockIdx.y;
  const int64_t rb = (int64_t)sp * p.chunk;
  int64_t re = rb + p.chunk;
  if (re > p.M) re = p.M;
  const int64_t ncol = p.N - c0;

  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const int64_t rstep = (int64_t)wpb * G;
#pragma unroll 4
  for (int64_t r = rb + (int64_t)w * G + grp; r < re; r += rstep) {
    const float xr = p.x[r];
    const f4 av = load4<VEC>(p.A + r * p.lda + c0, ncol);
    acc += av * xr;
  }
#pragma unroll
  for (int off = LPR; off < 64; off <<= 1) {
    acc[0] += __shfl_xor(acc[0], off);
    acc[1] += __shfl_xor(acc[1], off);
    acc[2] += __shfl_xor(acc[2], off);
    acc[3] += __shfl_xor(acc[3], off);
  }
  if (wpb > 1) {
    if (grp == 0) red[w][q] = acc;
    __syncthreads();
    if (w == 0 && grp == 0) {
      acc = red[0][q];
      for (int i = 1; i < wpb; ++i) acc += red[i][q];
    }
  }
  if (w == 0 && grp == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t col = c0 + t;
      if (col < p.N) {
        if (p.split == 1) {
          float *yp = p.y + col;
          float v = p.alpha * acc[t];
          if (p.read_c) v = p.alpha * acc[t] + p.beta * (*yp);
          *yp = v;
        } else {
          p.P[(int64_t)sp * p.N + col] = acc[t];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Streaming variants used when A, x are 16-byte aligned, lda % 4 == 0 and the
// contiguous extent is a multiple of 4: every lane's 4-float group is then
// wholly in range or wholly out.  Whole steps go through a guard-free
// pipeline with P steps of loads in flight in P named register buffers (the
// loop is unrolled by P, so nothing is copied and only the oldest step is
// waited for); A is loaded non-temporally (read once); the last partial
// step uses clamped addresses with zeroed values.
// ---------------------------------------------------------------------------
template <int LPR, int RB>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_dot_stream(GemvParams p) {
  constexpr int G = 64 / LPR;
  constexpr int P = RB <= 2 ? 4 : 2;
  constexpr int STEP = 4 * LPR;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)G * RB;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.split;
  const int sp = (int)(task - rb * p.split);
  if (rb >= nrb) return;  // with p.wsum the whole workgroup shares rb
  const int grp = lane / LPR, q = lane % LPR;

  const int64_t cb = (int64_t)sp * p.chunk;
  int64_t ce = cb + p.chunk;
  if (ce > p.N) ce = p.N;
  const int64_t nfull = ce > cb ? (ce - cb) / STEP : 0;

  const float *ap[RB];
  int64_t row[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    row[i] = rb * rows_per_wave + (int64_t)i * G + grp;
    const int64_t rc = row[i] < p.M ? row[i] : p.M - 1;
    ap[i] = p.A + rc * p.lda + cb + 4 * q;
  }
  const float *xp = p.x + cb + 4 * q;

  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.f;

  f4 bx[P], ba[P][RB];
  auto load = [&](int j, int64_t s) {
    bx[j] = *reinterpret_cast<const f4 *>(xp + s * STEP);
#pragma unroll
    for (int i = 0; i < RB; ++i)
      ba[j][i] = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(ap[i] + s * STEP));
  };
  auto comp = [&](int j) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      acc[i] += ba[j][i][0] * bx[j][0];
      acc[i] += ba[j][i][1] * bx[j][1];
      acc[i] += ba[j][i][2] * bx[j][2];
      acc[i] += ba[j][i][3] * bx[j][3];
    }
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
  {
    // last partial step: lanes past ce read column cb (in range) and zero it
    const int64_t c = cb + nfull * STEP + 4 * q;
    if (cb + nfull * STEP < ce) {
      const bool in = c < ce;
      const int64_t cc = in ? c : cb;
      f4 xv = *reinterpret_cast<const f4 *>(p.x + cc);
      if (!in) xv = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const int64_t rc = row[i] < p.M ? row[i] : p.M - 1;
        f4 av = *reinterpret_cast<const f4 *>(p.A + rc * p.lda + cc);
        if (!in) av = f4{0.f, 0.f, 0.f, 0.f};
        acc[i] += av[0] * xv[0];
        acc[i] += av[1] * xv[1];
        acc[i] += av[2] * xv[2];
        acc[i] += av[3] * xv[3];
      }
    }
  }

  dot_finish<LPR, RB>(p, acc, rb * rows_per_wave, sp);
}

template <int LPR>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_axpy_stream(GemvParams p) {
  constexpr int G = 64 / LPR;
  constexpr int CW = 4 * LPR;
  constexpr int P = 4;
  __shared__ f4 red[kGemvMaxWaves][LPR];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int wpb = blockDim.x >> 6;
  const int grp = lane / LPR, q = lane % LPR;
  const int64_t c0 = (int64_t)blockIdx.x * CW + 4 * q;
  const bool cin = c0 < p.N;          // N % 4 == 0: whole group in or out
  const int64_t cc = cin ? c0 : 0;
  const int sp = blockIdx.y;
  const int64_t rb = (int64_t)sp * p.chunk;
  int64_t re = rb + p.chunk;
  if (re > p.M) re = p.M;
  const int64_t rstep = (int64_t)wpb * G;
  const int64_t r0 = rb + (int64_t)w * G + grp;
  // whole steps: every lane group of the wave has its row inside [rb, re)
  const int64_t first_last = rb + (int64_t)w * G + (G - 1);
  const int64_t nfull = re > first_last ? (re - first_last - 1) / rstep + 1 : 0;

  f4 acc = {0.f, 0.f, 0.f, 0.f};
  const float *ap = p.A + r0 * p.lda + cc;
  const float *xp = p.x + r0;
  const int64_t astep = rstep * p.lda;
  f4 ba[P];
  float bx[P];
  auto load = [&](int j, int64_t s) {
    bx[j] = xp[s * rstep];
    ba[j] = __builtin_nontemporal_load(reinterpret_cast<const f4 *>(ap + s * astep));
  };
  auto comp = [&](int j) { acc += ba[j] * bx[j]; };
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
  {
    // rows after the whole steps (at most one partial step per lane group)
    const int64_t r = r0 + nfull * rstep;
    if (r < re) {
      const float xr = p.x[r];
      const f4 av = *reinterpret_cast<const f4 *>(p.A + r * p.lda + cc);
      acc += av * xr;
    }
  }
  if (!cin) acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int off = LPR; off < 64; off <<= 1) {
    acc[0] += __shfl_xor(acc[0], off);
    acc[1] += __shfl_xor(acc[1], off);
    acc[2] += __shfl_xor(acc[2], off);
    acc[3] += __shfl_xor(acc[3], off);
  }
  if (wpb > 1) {
    if (grp == 0) red[w][q] = acc;
    __syncthreads();
    if (w == 0 && grp == 0) {
      acc = red[0][q];
      for (int i = 1; i < wpb; ++i) acc += red[i][q];
    }
  }
  if (w == 0 && grp == 0 && cin) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t col = c0 + t;
      if (p.split == 1) {
        float *yp = p.y + col;
        float v = p.alpha * acc[t];
        if (p.read_c) v = p.alpha * acc[t] + p.beta * (*yp);
        *yp = v;
      } else {
        p.P[(int64_t)sp * p.N + col] = acc[t];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The same two pipelines for rows at any 4-byte alignment (odd lda, odd
// contiguous extent, unaligned A or x): 16-byte buffer loads need only dword
// alignment.  The descriptors are wave-uniform; every lane-dependent part of
// an address is a 32-bit lane offset, and a lane offset of kOOB (rows past
// the end, elements past the chunk or past N) reads zero without touching
// memory.  The host launches these only when every in-range lane offset is
// below 2^31 (gemv_buf_fits).
// ---------------------------------------------------------------------------
template <int LPR, int RB>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_dot_buf(GemvParams p) {
  constexpr int G = 64 / LPR;
  constexpr int P = RB <= 2 ? 4 : 2;
  constexpr int STEP = 4 * LPR;
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)G * RB;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.split;
  const int sp = (int)(task - rb * p.split);
  if (rb >= nrb) return;  // with p.wsum the whole workgroup shares rb
  const int grp = lane / LPR, q = lane % LPR;

  const int64_t cb = (int64_t)sp * p.chunk;
  int64_t ce = cb + p.chunk;
  if (ce > p.N) ce = p.N;
  const int64_t nfull = ce > cb ? (ce - cb) / STEP : 0;
  const int64_t row0 = rb * rows_per_wave;

  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.A + row0 * p.lda + cb, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rx = make_rsrc(p.x + cb, kRsrcMax);
  uint32_t aoff[RB];  // row row0 + iG + grp, column cb + 4q
#pragma unroll
  for (int i = 0; i < RB; ++i)
    aoff[i] = row0 + i * G + grp < p.M
                  ? (uint32_t)((((int64_t)i * G + grp) * p.lda + 4 * q) * 4) : kOOB;
  const uint32_t xoff = (uint32_t)(16 * q);

  float acc[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) acc[i] = 0.f;

  f4 bx[P], ba[P][RB];
  auto load = [&](int j, int64_t s) {
    const int32_t so = (int32_t)(s * STEP * 4);
    bx[j] = buf_load4<0>(rx, xoff, so);
#pragma unroll
    for (int i = 0; i < RB; ++i) ba[j][i] = buf_load4<kBufNT>(ra, aoff[i], so);
  };
  auto comp = [&](int j) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      acc[i] += ba[j][i][0] * bx[j][0];
      acc[i] += ba[j][i][1] * bx[j][1];
      acc[i] += ba[j][i][2] * bx[j][2];
      acc[i] += ba[j][i][3] * bx[j][3];
    }
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
        __builtin_amdgcn_sched_barrier(0);
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
  if (cb + nfull * STEP < ce) {
    // last partial step, element by element (past ce: out-of-range offset)
    const int64_t c = nfull * STEP + 4 * q;  // relative to cb
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool in = cb + c + e < ce;
      const float xv = buf_load1(rx, in ? (uint32_t)((c + e) * 4) : kOOB, 0);
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const uint32_t o = in && aoff[i] != kOOB ? aoff[i] + (uint32_t)((nfull * STEP + e) * 4) : kOOB;
        acc[i] += buf_load1(ra, o, 0) * xv;
      }
    }
  }

  dot_finish<LPR, RB>(p, acc, row0, sp);
}

template <int LPR>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_axpy_buf(GemvParams p) {
  constexpr int G = 64 / LPR;
  constexpr int CW = 4 * LPR;
  constexpr int P = 4;
  __shared__ f4 red[kGemvMaxWaves][LPR];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int grp = lane / LPR, q = lane % LPR;
  const int64_t cblk = (int64_t)blockIdx.x * CW;
  const int64_t c0 = cblk + 4 * q;
  const int sp = blockIdx.y;
  const int64_t rb = (int64_t)sp * p.chunk;
  int64_t re = rb + p.chunk;
  if (re > p.M) re = p.M;
  const int64_t rstep = (int64_t)wpb * G;
  const int64_t rw = rb + (int64_t)w * G;  // the wave's first row (uniform)
  const int64_t first_last = rw + (G - 1);
  const int64_t nfull = re > first_last ? (re - first_last - 1) / rstep + 1 : 0;

  // lane offsets of row rw + grp (+ s rstep through the descriptor base),
  // columns c0 .. c0 + 3; a group that crosses N loads element by element
  const int64_t ncol = p.N - c0;
  const bool whole = ncol >= 4;
  const uint32_t aoff = (uint32_t)(((int64_t)grp * p.lda + 4 * q) * 4);
  const uint32_t xoff = (uint32_t)(4 * grp);
  const __amdgpu_buffer_rsrc_t rx = make_rsrc(p.x + rw, kRsrcMax);

  f4 acc = {0.f, 0.f, 0.f, 0.f};
  f4 ba[P];
  float bx[P];
  auto loadrow = [&](int64_t s, bool in) -> f4 {
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(p.A + (rw + s * rstep) * p.lda + cblk, kRsrcMax);
    if (whole) return buf_load4<kBufNT>(rs, in ? aoff : kOOB, 0);
    f4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      v[e] = buf_load1(rs, in && ncol > e ? aoff + 4 * e : kOOB, 0);
    return v;
  };
  auto load = [&](int j, int64_t s) {
    bx[j] = buf_load1(rx, xoff, (int32_t)(s * rstep * 4));
    ba[j] = loadrow(s, true);
  };
  auto comp = [&](int j) { acc += ba[j] * bx[j]; };
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
        __builtin_amdgcn_sched_barrier(0);
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
  {
    // rows after the whole steps (at most one partial step per lane group)
    const bool in = rw + nfull * rstep + grp < re;
    if (rw + nfull * rstep < re) {
      const float xr = buf_load1(rx, in ? xoff : kOOB, (int32_t)(nfull * rstep * 4));
      acc += loadrow(nfull, in) * xr;
    }
  }
#pragma unroll
  for (int off = LPR; off < 64; off <<= 1) {
    acc[0] += __shfl_xor(acc[0], off);
    acc[1] += __shfl_xor(acc[1], off);
    acc[2] += __shfl_xor(acc[2], off);
    acc[3] += __shfl_xor(acc[3], off);
  }
  if (wpb > 1) {
    if (grp == 0) red[w][q] = acc;
    __syncthreads();
    if (w == 0 && grp == 0) {
      acc = red[0][q];
      for (int i = 1; i < wpb; ++i) acc += red[i][q];
    }
  }
  if (w == 0 && grp == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int64_t col = c0 + t;
      if (col < p.N) {
        if (p.split == 1) {
          float *yp = p.y + col;
          float v = p.alpha * acc[t];
          if (p.read_c) v = p.alpha * acc[t] + p.beta * (*yp);
          *yp = v;
        } else {
          p.P[(int64_t)sp * p.N + col] = acc[t];
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Several right-hand sides (skinny SGEMM with NV = 2, 4 or 8 columns, A
// contiguous along K): C(m, n) = sum_k A(m, k) B(k, n).  The DOT streaming
// structure with NV accumulators per row: lane groups stream rows of A with
// 128-bit buffer loads (any dword alignment); per 4-k step a lane loads
// B(k .. k+3, 0 .. NV-1) once for its RB rows -- NV 16-byte loads along k
// when B is K-contiguous (BVEC), else four rows of NV contiguous floats --
// and does 4 NV FMAs per row.  That is NV FMAs per element of A, within the
// VALU rate at HBM speed up to NV = 8, with no LDS staging or matrix-core
// padding of the columns (a 16-column MFMA tile for 2 columns).  Loads of
// steps past the chunk go through a size-0 descriptor (zeros, no traffic),
// so the pipeline needs no guarded drain; a partial last step is read
// element by element.  K splits follow the GEMV DOT rules (wsum: the waves
// of a workgroup add their chunks in LDS in wave order).
// ---------------------------------------------------------------------------
template <int LPR, int RB, int NV, bool BVEC>
__global__ __launch_bounds__(64 * kGemvMaxWaves) void k_gemv_multi(GemvNParams p) {
  constexpr int G = 64 / LPR;
  constexpr int PS = RB * NV <= 8 ? 3 : 2;  // steps of loads in flight
  constexpr int STEP = 4 * LPR;
  __shared__ float red[kGemvMaxWaves * RB * G * NV];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + w;
  const int64_t rows_per_wave = (int64_t)G * RB;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.split;
  const int sp = (int)(task - rb * p.split);
  if (rb >= nrb) return;  // with p.wsum the whole workgroup shares rb
  const int grp = lane / LPR, q = lane % LPR;

  const int64_t cb = (int64_t)sp * p.chunk;
  int64_t ce = cb + p.chunk;
  if (ce > p.K) ce = p.K;
  const int64_t nfull = ce > cb ? (ce - cb) / STEP : 0;
  const int64_t row0 = rb * rows_per_wave;

  const __amdgpu_buffer_rsrc_t ra = make_rsrc(p.A + row0 * p.lda + cb, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rbs = make_rsrc(p.B + cb * p.sBk, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rnull = make_rsrc(p.A, 0);
  uint32_t aoff[RB];  // row row0 + iG + grp, k = cb + 4q
#pragma unroll
  for (int i = 0; i < RB; ++i)
    aoff[i] = row0 + i * G + grp < p.M
                  ? (uint32_t)((((int64_t)i * G + grp) * p.lda + 4 * q) * 4) : kOOB;
  // B(cb + 4q + e, n): BVEC column n at boff[n] (e along the 16 bytes),
  // else row e at boff[e] (n along the row)
  uint32_t boff[BVEC ? NV : 4];
#pragma unroll
  for (int i = 0; i < (BVEC ? NV : 4); ++i)
    boff[i] = BVEC ? (uint32_t)((i * p.sBn + 4 * q) * 4) : (uint32_t)((4 * q + i) * p.sBk * 4);

  float acc[RB][NV];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int n = 0; n < NV; ++n) acc[i][n] = 0.f;
  const bool dense2 = !BVEC && NV == 2 && p.sBk == 2 && p.sBn == 1;

  f4 ba[PS][RB];
  float bx[PS][4][NV];
  auto load = [&](int j, int64_t s) {
    const bool live = s < nfull;
    const __amdgpu_buffer_rsrc_t rs_a = live ? ra : rnull;
    const __amdgpu_buffer_rsrc_t rs_b = live ? rbs : rnull;
    const int32_t so = (int32_t)(s * STEP * 4);
#pragma unroll
    for (int i = 0; i < RB; ++i) ba[j][i] = buf_load4<kBufNT>(rs_a, aoff[i], so);
    if (BVEC) {
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        const f4 t = buf_load4<0>(rs_b, boff[n], so);
#pragma unroll
        for (int e = 0; e < 4; ++e) bx[j][e][n] = t[e];
      }
    } else {
      const int32_t sob = (int32_t)(s * STEP * p.sBk * 4);
      if (NV == 2 && dense2) {
        // rows of 2 floats back to back: the 4 rows are 32 contiguous bytes
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const f4 t = buf_load4<0>(rs_b, boff[0] + 16 * h, sob);
          bx[j][2 * h][0] = t[0];
          bx[j][2 * h][1] = t[1];
          bx[j][2 * h + 1][0] = t[2];
          bx[j][2 * h + 1][1] = t[3];
        }
      } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (NV == 2) {
          const f2 t = buf_load2(rs_b, boff[e], sob);
          bx[j][e][0] = t[0];
          bx[j][e][1] = t[1];
        } else {
#pragma unroll
          for (int h = 0; h < NV / 4; ++h) {
            const f4 t = buf_load4<0>(rs_b, boff[e] + 16 * h, sob);
#pragma unroll
            for (int n = 0; n < 4; ++n) bx[j][e][4 * h + n] = t[n];
          }
        }
      }
      }
    }
  };
  auto comp = [&](int j) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int n = 0; n < NV; ++n) acc[i][n] += ba[j][i][e] * bx[j][e][n];
  };

#pragma unroll
  for (int j = 0; j < PS; ++j) load(j, j);
  __builtin_amdgcn_sched_barrier(0);
  for (int64_t s = 0; s < nfull; s += PS) {
#pragma unroll
    for (int j = 0; j < PS; ++j) {
      if (s + j < nfull) comp(j);
      load(j, s + PS + j);
      __builtin_amdgcn_sched_barrier(0);  // keep each step's loads right after its compute
    }
  }
  if (cb + nfull * STEP < ce) {
    // last partial step, element by element (past ce: out-of-range offset)
    const int64_t c = nfull * STEP + 4 * q;  // relative to cb
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const bool in = cb + c + e < ce;
      float xv[NV];
#pragma unroll
      for (int n = 0; n < NV; ++n)
        xv[n] = buf_load1(rbs, in ? (uint32_t)(((c + e) * p.sBk + n * p.sBn) * 4) : kOOB, 0);
#pragma unroll
      for (int i = 0; i < RB; ++i) {
        const uint32_t o = in && aoff[i] != kOOB ? aoff[i] + (uint32_t)((nfull * STEP + e) * 4) : kOOB;
        const float av = buf_load1(ra, o, 0);
#pragma unroll
        for (int n = 0; n < NV; ++n) acc[i][n] += av * xv[n];
      }
    }
  }

  // lane groups' sums; with p.wsum the waves of the workgroup meet in LDS
  // and wave 0 adds them in wave order
  float sm[RB][NV];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int n = 0; n < NV; ++n) sm[i][n] = group_sum<LPR>(acc[i][n]);
  int slot = sp, nslot = p.split;
  if (p.wsum) {
    if (q == 0) {
#pragma unroll
      for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int n = 0; n < NV; ++n) red[((w * RB + i) * G + grp) * NV + n] = sm[i][n];
    }
    __syncthreads();
    if (w != 0) return;
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        float t = red[(i * G + grp) * NV + n];
        for (int ww = 1; ww < wpb; ++ww) t += red[((ww * RB + i) * G + grp) * NV + n];
        sm[i][n] = t;
      }
    slot = sp / wpb;
    nslot = p.split / wpb;
  }
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int64_t row = row0 + (int64_t)i * G + grp;
    if (q == 0 && row < p.M) {
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        if (nslot == 1) {
          float *cp = p.C + row * p.sCm + n * p.sCn;
          float v = p.alpha * sm[i][n];
          if (p.read_c) v = p.alpha * sm[i][n] + p.beta * (*cp);
          *cp = v;
        } else {
          p.P[((int64_t)slot * p.M + row) * NV + n] = sm[i][n];
        }
      }
    }
  }
}

template <int LPR, int RB, int NV, bool BVEC>
hipError_t run_multi(int wpb, const GemvNParams &p, hipStream_t st) {
  const int64_t rows = (int64_t)(64 / LPR) * RB;
  const int64_t tasks = ((p.M + rows - 1) / rows) * p.split;
  const int64_t nblk = (tasks + wpb - 1) / wpb;
  if (nblk > 0x7fffffff) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL((k_gemv_multi<LPR, RB, NV, BVEC>), dim3((unsigned)nblk), dim3(64 * wpb), 0,
                     st, p);
  return hipGetLastError();
}

typedef hipError_t (*MultiFn)(int, const GemvNParams &, hipStream_t);

template <int LPR, int RB, int NV>
MultiFn multi_fn(bool bvec) {
  return bvec ? run_multi<LPR, RB, NV, true> : run_multi<LPR, RB, NV, false>;
}

template <int LPR, int RB>
MultiFn multi_nv(int nv, bool bvec) {
  switch (nv) {
    case 2: return multi_fn<LPR, RB, 2>(bvec);
    case 4: return multi_fn<LPR, RB, 4>(bvec);
    case 8: return multi_fn<LPR, RB, 8>(bvec);
    default: return nullptr;
  }
}

template <int LPR>
MultiFn multi_rb(int rb, int nv, bool bvec) {
  switch (rb) {
    case 1: return multi_nv<LPR, 1>(nv, bvec);
    case 2: return multi_nv<LPR, 2>(nv, bvec);
    case 4: return multi_nv<LPR, 4>(nv, bvec);
    default: return nullptr;
  }
}

MultiFn find_multi(int lpr, int rb, int nv, bool bvec) {
  switch (lpr) {
    case 16: return multi_rb<16>(rb, nv, bvec);
    case 32: return multi_rb<32>(rb, nv, bvec);
    case 64: return multi_rb<64>(rb, nv, bvec);
    default: return nullptr;
  }
}

// Every in-range lane offset of the buffer kernels is below 2^31.
bool gemv_buf_fits(int family, int lpr, int rb, const GemvParams &p) {
  const int64_t G = 64 / lpr;
  const int64_t lim = (int64_t)1 << 31;
  if (family == 1)  // rows of a wave x lda + the chunk
    return ((G * rb) * p.lda + p.chunk + 4) * 4 < lim;
  return (G * p.lda + 4 * lpr + 4) * 4 < lim && (p.M + 4) * 4 < lim;
}

typedef hipError_t (*RunFn)(int, const GemvParams &, hipStream_t);

template <int LPR, int RB, bool VEC>
hipError_t run_dot(int wpb, const GemvParams &p, hipStream_t st) {
  const int64_t rows = (int64_t)(64 / LPR) * RB;
  const int64_t tasks = ((p.M + rows - 1) / rows) * p.split;
  const int64_t nblk = (tasks + wpb - 1) / wpb;
  if (nblk > 0x7fffffff) return hipErrorInvalidConfiguration;
  if (VEC)
    hipLaunchKernelGGL((k_gemv_dot_stream<LPR, RB>), dim3((unsigned)nblk),
                       dim3(64 * wpb), 0, st, p);
  else if (gemv_buf_fits(1, LPR, RB, p))
    hipLaunchKernelGGL((k_gemv_dot_buf<LPR, RB>), dim3((unsigned)nblk),
                       dim3(64 * wpb), 0, st, p);
  else
    hipLaunchKernelGGL((k_gemv_dot<LPR, RB, false>), dim3((unsigned)nblk),
                       dim3(64 * wpb), 0, st, p);
  return hipGetLastError();
}

template <int LPR, bool VEC>
hipError_t run_axpy(int wpb, const GemvParams &p, hipStream_t st) {
  const int64_t ncb = (p.N + 4 * LPR - 1) / (4 * LPR);
  if (ncb > 0x7fffffff || p.split > 65535) return hipErrorInvalidConfiguration;
  if (VEC)
    hipLaunchKernelGGL((k_gemv_axpy_stream<LPR>), dim3((unsigned)ncb, (unsigned)p.split),
                       dim3(64 * wpb), 0, st, p);
  else if (gemv_buf_fits(2, LPR, 1, p))
    hipLaunchKernelGGL((k_gemv_axpy_buf<LPR>), dim3((unsigned)ncb, (unsigned)p.split),
                       dim3(64 * wpb), 0, st, p);
  else
    hipLaunchKernelGGL((k_gemv_axpy<LPR, false>), dim3((unsigned)ncb, (unsigned)p.split),
                       dim3(64 * wpb), 0, st, p);
  return hipGetLastError();
}

template <int LPR>
RunFn dot_rb(int rb, bool vec) {
  switch (rb) {
    case 1: return vec ? run_dot<LPR, 1, true> : run_dot<LPR, 1, false>;
    case 2: return vec ? run_dot<LPR, 2, true> : run_dot<LPR, 2, false>;
    case 4: return vec ? run_dot<LPR, 4, true> : run_dot<LPR, 4, false>;
    case 8: return vec ? run_dot<LPR, 8, true> : run_dot<LPR, 8, false>;
    default: return nullptr;
  }
}

RunFn find_gemv(int family, int lpr, int rb, bool vec) {
  if (family == 1) {  // DOT
    switch (lpr) {
      case 4: return dot_rb<4>(rb, vec);
      case 8: return dot_rb<8>(rb, vec);
      case 16: return dot_rb<16>(rb, vec);
      case 32: return dot_rb<32>(rb, vec);
      case 64: return dot_rb<64>(rb, vec);
      default: return nullptr;
    }
  } else if (family == 2) {  // AXPY; rb is unused (fixed unroll)
    if (rb != 1) return nullptr;
    switch (lpr) {
      case 4: return vec ? run_axpy<4, true> : run_axpy<4, false>;
      case 8: return vec ? run_axpy<8, true> : run_axpy<8, false>;
      case 16: return vec ? run_axpy<16, true> : run_axpy<16, false>;
      case 32: return vec ? run_axpy<32, true> : run_axpy<32, false>;
      case 64: return vec ? run_axpy<64, true> : run_axpy<64, false>;
      default: return nullptr;
    }
  }
  return nullptr;
}

}  // namespace

bool gemv_multi_has_variant(int lpr, int rb, int nv) {
  return find_multi(lpr, rb, nv, true) != nullptr;
}

bool gemv_multi_offsets_fit(int lpr, int rb, int nv, bool bvec, const GemvNParams &p) {
  const double lim = 2147483647.0;
  const double a = ((double)(64 / lpr) * rb * (double)p.lda + (double)p.chunk + 4.0) * 4.0;
  const double b = bvec ? ((double)nv * (double)p.sBn + (double)p.chunk + 4.0) * 4.0
                        : (((double)p.chunk + 4.0) * (double)p.sBk + nv) * 4.0;
  return a < lim && b < lim;
}

hipError_t launch_gemv_multi(int lpr, int rb, int nv, bool bvec, int wpb, const GemvNParams &p,
                             hipStream_t st) {
  MultiFn fn = find_multi(lpr, rb, nv, bvec);
  if (!fn || !gemv_wpb_ok(1, wpb) || p.split < 1 || (p.wsum && p.split % wpb != 0) ||
      !gemv_multi_offsets_fit(lpr, rb, nv, bvec, p))
    return hipErrorInvalidValue;
  return fn(wpb, p, st);
}

bool gemv_wpb_ok(int, int wpb) {
  return wpb == 1 || wpb == 2 || wpb == 4 || wpb == kGemvMaxWaves;
}

bool gemv_has_variant(int family, int lpr, int rb) {
  return find_gemv(family, lpr, rb, true) != nullptr;
}

hipError_t launch_gemv(int family, int lpr, int rb, bool vec, int wpb,
                       const GemvParams &p, hipStream_t st) {
  RunFn fn = find_gemv(family, lpr, rb, vec);
  if (!fn || !gemv_wpb_ok(family, wpb) || p.split < 1 || (p.wsum && p.split % wpb != 0))
    return hipErrorInvalidValue;
  return fn(wpb, p, st);
}

}  // namespace hml
