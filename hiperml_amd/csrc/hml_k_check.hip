// hml_k_check.hip -- FP64 check kernel (the GPU half of the FP64 oracle).
//
// R(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C0(m,n) in double for any
// strided layout, compared in the same pass with the FP32 result C(m,n).
// Each block reduces its 16x16 tile to six statistics (see hml_ext.h:
// hmlSgemmCheckF64); the host adds the per-block values in a fixed order.
// The reference has no such check (its apps compare a double checksum of C,
// apps/gemm/hml_sgemm_main.cu:497-511); this replaces that checksum test with
// a per-element one.
#include "hml_internal.h"

namespace hml {
namespace {

__global__ __launch_bounds__(256) void k_check_f64(CheckParams p) {
  __shared__ double As[16][17];
  __shared__ double Bs[16][17];
  __shared__ double red[6][256];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t row = (int64_t)blockIdx.x * 16 + ty;
  const int64_t col = (int64_t)blockIdx.y * 16 + tx;
  double s = 0.0, sa = 0.0;
  for (int64_t k0 = 0; k0 < p.K; k0 += 16) {
    {
      const int64_t ar = (int64_t)blockIdx.x * 16 + ty, ak = k0 + tx;
      As[ty][tx] = (ar < p.M && ak < p.K) ? (double)p.A[ar * p.sAm + ak * p.sAk] : 0.0;
      const int64_t bk = k0 + ty, bn = (int64_t)blockIdx.y * 16 + tx;
      Bs[ty][tx] = (bk < p.K && bn < p.N) ? (double)p.B[bk * p.sBk + bn * p.sBn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const double a = As[ty][kk], b = Bs[kk][tx];
      s = fma(a, b, s);
      sa = fma(fabs(a), fabs(b), sa);
    }
    __syncthreads();
  }
  double e2 = 0.0, r2 = 0.0, b2 = 0.0, emax = 0.0, erel = 0.0, nonfin = 0.0;
  if (row < p.M && col < p.N) {
    double c0 = 0.0;
    if (p.C0) c0 = (double)p.C0[row * p.sCm + col * p.sCn];
    const double R = p.alpha * s + (p.beta != 0.0 ? p.beta * c0 : 0.0);
    const double bound = fabs(p.alpha) * sa + fabs(p.beta) * fabs(c0);
    const float cf = p.C[row * p.sCm + col * p.sCn];
    const double c = (double)cf;
    if (!isfinite(cf) && isfinite(R)) {
      nonfin = 1.0;
    } else {
      const double e = fabs(c - R);
      e2 = e * e;
      r2 = R * R;
      b2 = bound * bound;
      emax = e;
      erel = bound > 0.0 ? e / bound : (e > 0.0 ? 1.0e300 : 0.0);
    }
  }
  const int t = threadIdx.x;
  red[0][t] = e2; red[1][t] = r2; red[2][t] = b2;
  red[3][t] = emax; red[4][t] = erel; red[5][t] = nonfin;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
      red[2][t] += red[2][t + h];
      red[3][t] = fmax(red[3][t], red[3][t + h]);
      red[4][t] = fmax(red[4][t], red[4][t + h]);
      red[5][t] += red[5][t + h];
    }
    __syncthreads();
  }
  if (t < 6) {
    const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    p.stats[b * 6 + t] = red[t][0];
  }
}

}  // namespace

hipError_t launch_check_f64(const CheckParams &p, int64_t *nblocks,
                            hipStream_t st) {
  const int64_t gm = (p.M + 15) / 16, gn = (p.N + 15) / 16;
  *nblocks = gm * gn;
  if (gm <= 0 || gn <= 0) return hipSuccess;
  if (gm > 0x7fffffff || gn > 65535) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(k_check_f64, dim3((unsigned)gm, (unsigned)gn), dim3(256),
                     0, st, p);
  return hipGetLastError();
}

}  // namespace hml

// ---------------------------------------------------------------------------
// Device-side synthetic inputs and output comparison (autotuner, benches).
// ---------------------------------------------------------------------------
namespace hml {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// dst[i] = lo + (hi - lo) * u_i, u_i in [0,1) from the top 24 bits of
// splitmix64(seed * 2^32 + i): a counter-based stream, identical for any
// launch geometry.
__global__ void k_fill_uniform(float *dst, size_t n, uint64_t seed, float lo,
                               float hi) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    const uint64_t z = splitmix64((seed << 32) + i);
    const float u = (float)(z >> 40) * (1.0f / 16777216.0f);
    dst[i] = lo + (hi - lo) * u;
  }
}

__global__ __launch_bounds__(256) void k_compare(const float *C,
                                                 const float *R, int64_t M,
                                                 int64_t N, int64_t sCm,
                                                 int64_t sCn, double *stats) {
  __shared__ double red[5][256];
  double e2 = 0.0, r2 = 0.0, emax = 0.0, rmax = 0.0, nonfin = 0.0;
  const int64_t total = M * N;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t m = idx / N, n = idx - m * N;
    const float c = C[m * sCm + n * sCn];
    const float r = R[m * sCm + n * sCn];
    if (!isfinite(c) && isfinite(r)) { nonfin += 1.0; continue; }
    const double e = fabs((double)c - (double)r);
    e2 += e * e;
    r2 += (double)r * (double)r;
    emax = fmax(emax, e);
    rmax = fmax(rmax, fabs((double)r));
  }
  const int t = threadIdx.x;
  red[0][t] = e2; red[1][t] = r2; red[2][t] = emax; red[3][t] = rmax; red[4][t] = nonfin;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      red[0][t] += red[0][t + h];
      red[1][t] += red[1][t + h];
      red[2][t] = fmax(red[2][t], red[2][t + h]);
      red[3][t] = fmax(red[3][t], red[3][t + h]);
      red[4][t] += red[4][t + h];
    }
    __syncthreads();
  }
  if (t < 5) stats[(int64_t)blockIdx.x * 5 + t] = red[t][0];
}

// Read-only stream: each wave reads 8 KB contiguous pieces (8 loads of
// 1 KB, 16 bytes per lane, all in flight together), pieces grid-strided, and
// folds them with XOR; the fold is stored only in the never-taken case that
// it equals a fixed odd constant, so the kernel moves no bytes except the
// reads.
typedef uint32_t probe_u4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_read_probe(const probe_u4 *__restrict__ src,
                                                    size_t n16, uint32_t *sink) {
  const int lane = threadIdx.x & 63;
  const size_t nw = (size_t)gridDim.x * (blockDim.x >> 6);
  const size_t gw = (size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  probe_u4 acc = {0u, 0u, 0u, 0u};
  for (size_t base = gw * 512; base < n16; base += nw * 512) {
    probe_u4 v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const size_t i = base + 64 * j + lane;
      v[j] = i < n16 ? __builtin_nontemporal_load(src + i) : probe_u4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) acc ^= v[j];
  }
  const uint32_t f = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (f == 0x9E3779B9u) sink[0] = f;
}

}  // namespace

hipError_t launch_read_probe(const void *src, size_t bytes, uint32_t *sink, hipStream_t st) {
  const size_t n16 = bytes / 16;
  if (n16 == 0) return hipSuccess;
  // 8 workgroups of 4 waves per CU on 256 CUs
  hipLaunchKernelGGL(k_read_probe, dim3(2048), dim3(256), 0, st, (const probe_u4 *)src, n16,
                     sink);
  return hipGetLastError();
}

hipError_t launch_fill_uniform(float *dst, size_t n, uint64_t seed, float lo,
                               float hi, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t nblk = (n + 255) / 256;
  if (nblk > 65536) nblk = 65536;
  hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)nblk), dim3(256), 0, st,
                     dst, n, seed, lo, hi);
  return hipGetLastError();
}

hipError_t launch_compare(const float *C, const float *R, int64_t M, int64_t N,
                          int64_t sCm, int64_t sCn, double *stats, int nblocks,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_compare, dim3((unsigned)nblocks), dim3(256), 0, st, C,
                     R, M, N, sCm, sCn, stats);
  return hipGetLastError();
}

}  // namespace hml
