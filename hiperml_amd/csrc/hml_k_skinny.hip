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

// ---------------------------------------------------------------------------
// DOT form: A(m,k) = A[m*sAm + k].  Per k-step of 32, lane (g = l/16,
// r = l%16) of row tile t loads A(row_t + r, k0 + 8g .. k0 + 8g + 7) as two
// 128-bit loads; MFMA number 4v+j consumes element j of load v, so the
// shared K permutation is k = k0 + 8g + 4v + j.  The packed B fragment of
// (step s, column tile c, half v) holds, for lane (g, r),
// B(32s + 8g + 4v + j, 16c + r) for j = 0..3.
// ---------------------------------------------------------------------------
template <int RT, int CT, bool VEC>
__global__ __launch_bounds__(256) void k_skinny_dot(SkinnyParams p) {
  const int lane = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  const int64_t task = (int64_t)blockIdx.x * wpb + (threadIdx.x >> 6);
  const int64_t rows_per_wave = (int64_t)kDotRowsPerTile * RT;
  const int64_t nrb = (p.M + rows_per_wave - 1) / rows_per_wave;
  const int64_t rb = task / p.ksplit;
  const int ks = (int)(task - rb * p.ksplit);
  if (rb >= nrb) return;  // whole wave leaves; the kernel has no barriers
  const int ct0 = blockIdx.y * CT;
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

  const f4 *Bp = reinterpret_cast<const f4 *>(p.Bp);
  f4 acc[RT][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = f4{0.f, 0.f, 0.f, 0.f};

  f4 a[RT][2], b[CT][2];
  auto load_step = [&](int64_t k0, f4 (&aa)[RT][2], f4 (&bb)[CT][2]) {
    const int64_t kk = k0 + 8 * g;
#pragma unroll
    for (int t = 0; t < RT; ++t) {
      aa[t][0] = load4<VEC>(arow[t] + kk, ke - kk);
      aa[t][1] = load4<VEC>(arow[t] + kk + 4, ke - kk - 4);
    }
    const int64_t s = k0 / kDotKStep;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      const int64_t base = ((s * p.nct_pad + ct0 + c) * 2) * 64 + lane;
      bb[c][0] = Bp[base];
      bb[c][1] = Bp[base + 64];
    }
  };

  if (kb < ke) load_step(kb, a, b);
  for (int64_t k0 = kb; k0 < ke; k0 += kDotKStep) {
    f4 an[RT][2], bn[CT][2];
    const bool more = k0 + kDotKStep < ke;
    if (more) load_step(k0 + kDotKStep, an, bn);
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < RT; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[t][c] = mfma4(a[t][v][j], b[c][v][j], acc[t][c]);
    if (more) {
#pragma unroll
      for (int t = 0; t < RT; ++t) { a[t][0] = an[t][0]; a[t][1] = an[t][1]; }
#pragma unroll
      for (int c = 0; c < CT; ++c) { b[c][0] = bn[c][0]; b[c][1] = bn[c][1]; }
    }
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

// ---------------------------------------------------------------------------
// AXPY form: A(m,k) = A[m + k*sAk].  A row tile is 64 rows: lane (g, r)
// loads A(m0 + 4r .. m0 + 4r + 3, k) for k = k0 + 4u + g (u = 0..3 per
// k-step of 16).  MFMA j (j = 0..3) takes element j, so its 16 "rows" are
// the interleaved rows m0 + 4r + j; its B operand B(k0 + 4u + g, 16c + r)
// is shared by the 4 MFMAs.  Packed B fragment (step s, tile c) holds, for
// lane (g, r), B(16s + 4u + g, 16c + r) for u = 0..3.
// ---------------------------------------------------------------------------
template <int RT, int CT, bool VEC>
__global__ __launch_bounds__(256) void k_skinny_axpy(SkinnyParams p) {
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

  int64_t m4[RT], mvalid[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    m4[t] = rb * rows_per_wave + t * 64 + 4 * r;
    mvalid[t] = p.M - m4[t];
    if (mvalid[t] < 0) { mvalid[t] = 0; m4[t] = 0; }
  }
  const int64_t kb = (int64_t)ks * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;

  const f4 *Bp = reinterpret_cast<const f4 *>(p.Bp);
  f4 acc[RT][4][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][j][c] = f4{0.f, 0.f, 0.f, 0.f};

  f4 a[RT][4], b[CT];
  auto load_step = [&](int64_t k0, f4 (&aa)[RT][4], f4 (&bb)[CT]) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = k0 + 4 * u + g;
      const bool kin = k < ke;
#pragma unroll
      for (int t = 0; t < RT; ++t)
        aa[t][u] = load4<VEC>(p.A + m4[t] + (kin ? k : 0) * p.sAk,
                              kin ? mvalid[t] : 0);
    }
    const int64_t s = k0 / kAxpyKStep;
#pragma unroll
    for (int c = 0; c < CT; ++c)
      bb[c] = Bp[(s * p.nct_pad + ct0 + c) * 64 + lane];
  };

  if (kb < ke) load_step(kb, a, b);
  for (int64_t k0 = kb; k0 < ke; k0 += kAxpyKStep) {
    f4 an[RT][4], bn[CT];
    const bool more = k0 + kAxpyKStep < ke;
    if (more) load_step(k0 + kAxpyKStep, an, bn);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[t][j][c] = mfma4(a[t][u][j], b[c][u], acc[t][j][c]);
    if (more) {
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) a[t][u] = an[t][u];
#pragma unroll
      for (int c = 0; c < CT; ++c) b[c] = bn[c];
    }
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
__global__ void k_pack_b(const float *__restrict__ B, int64_t sBk, int64_t sBn,
                         int64_t K, int64_t N, int form, int64_t total,
                         int nct_pad, f4 *__restrict__ Bp) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
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
    kbase = s * kDotKStep + 8 * g + 4 * half;
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
  Bp[idx] = v;
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
      case 2: HML_CT_CASES(axpy_fn, 2)
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
