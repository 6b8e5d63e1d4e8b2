// hml_k_skinny.hip -- skinny SGEMM on gfx950 FP32 matrix cores.
//
// Replaces the reference's register-blocked CUDA tile kernels
// hmlSgemmVarKNN / hmlSgemmConstKNN (include/hml_sgemm_kernel_template.h:
// 19-191, 229-333).  Same operation C = alpha*A*B + beta*C; different
// machine model:
//
//  * FP32 products run on the matrix cores (2x the non-packed VALU rate).
//    v_mfma_f32_16x16x4f32: lane l supplies A[m=l%16][k=l/16] and
//    B[k=l/16][n=l%16]; the 16x16 result sits in 4 VGPRs, lane l holding
//    D[4*(l/16)+i][l%16].  The reduction index k is a sum, so any
//    permutation of K that A and B share is legal; each kernel picks the
//    permutation that makes its A loads contiguous.
//  * The streamed operand A (the big one) is read exactly once from HBM.
//    DOT form (A contiguous along K): k_skinny_dot loads whole 256-byte row
//    pieces, transposes them through a per-wave LDS tile and reads B in
//    place; k_skinny_dotb shares each B stage between the 4 waves of a
//    workgroup through LDS; k_grouped_dot runs many problems in one
//    persistent launch.  AXPY form (A contiguous along M): k_skinny_axpy
//    reads A in operand order (4 columns x 256 bytes per load) and B in
//    place.
//  * K is split across waves (inside a workgroup, summed in LDS) and, when
//    rows alone cannot fill 256 CUs, across workgroups, whose partial tiles
//    are summed in a fixed order by a second kernel (deterministic, no
//    atomics).
#include <algorithm>
#include <mutex>

#include "hml_internal.h"
#include "hml_kdev.h"

namespace hml {
namespace {

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// The skinny kernels stream K in steps (stages).  Steps that lie wholly
// inside the wave's K range go through a pipeline that keeps PS steps of
// loads in flight in PS named register sets (the loop is unrolled by PS, so
// no set is ever copied and the compiler waits only for the oldest step);
// loads of steps past the range go through a descriptor of size 0 (zeros,
// no memory traffic), and a partial last step (K not a multiple of the
// step) is loaded with bounds checks afterwards.

// ---------------------------------------------------------------------------
// DOT form, one problem: k_skinny_dot.
//
// A wave computes rows [row0, row0 + 16 RT) x column tiles [ct0, ct0 + CT)
// over its own K chunk; the kw waves of a workgroup take consecutive chunks
// of the same tile and sum their results through LDS at the end, in wave
// order, so a K split needs no second kernel unless it spans workgroups.
//
// A streams in stages of 64 k.  The global loads of a stage are
// row-contiguous: load instruction q moves rows row0 + 4q + h (h = l/16) at
// k = 4(l%16) .. +3, i.e. 4 rows x 256 contiguous bytes -- whole 128-byte
// lines -- where the MFMA operand order itself would fetch 64 bytes of each
// of 16 rows per instruction (half lines, twice the line requests per byte).
// The stage goes into the wave's own LDS tile (16 RT rows x 64 floats, the
// 16-byte chunk kc of row x stored at chunk kc ^ (x % 16), which spreads
// both the row-contiguous writes and the operand-order reads over all 64
// banks) and comes back in operand order: lane (g, r) takes
// A(row r, k0 + 16v + 4g .. +3), v = 0, 1, per k-step of 32.  The tile is
// private to the wave, so only the wave's own LDS ordering is needed, not a
// barrier.
//
// B(k, n) is read in place with its own strides: lane
// (g, r) of column tile c needs B(k0 + 16v + 4g + j, 16c + r), j = 0..3 --
// one 16-byte load when B is K-contiguous and aligned (BVEC), four scalar
// loads otherwise.  Columns past N read column N - 1: they feed only output
// columns that are never stored.  K past the chunk end reads as zero (A and
// B), so no NaN from memory outside the operands can reach a sum.
//
// The A and B loads of a stage are issued together, PS stages ahead, into
// PS register sets: all vector loads retire through one in-order counter,
// so waiting for a stage's B must not also wait for later stages' A.  Tiles
// of 4 column tiles (FP32-bound) instead load A straight into PS LDS tiles
// (LDS DMA, see dot_tile_lds), which frees the A registers.
// ---------------------------------------------------------------------------
// pipeline depth (stages of A and B loads in flight) and waves per
// workgroup of each tile shape, from its register budget.  Two stages per
// wave measured as fast as three or four on MI355X (8192 x 8192 A, N <= 32)
// and leave registers for more resident waves.
__host__ __device__ constexpr int dot_stages(int, int) { return 2; }
__host__ __device__ constexpr int dot_max_waves(int rt, int ct) {
  return (rt * ct >= 8 || rt == 4) ? 4 : 8;
}
// floats of one A tile (16 rt rows x one 64-k stage)
__host__ __device__ constexpr int dot_tile_floats(int rt) { return rt * 16 * kDotStageK; }
// k_skinny_dot stages A through LDS DMA for 4-column-tile (FP32-bound)
// shapes: measured on 8192 x {64, 128} x 8192, 2-5% faster than the register
// path; with fewer columns the extra LDS tile costs more occupancy than the
// saved registers buy (up to 30% slower for N = 2).
__host__ __device__ constexpr bool dot_dma(int, int ct) { return ct >= 3; }
// floats of LDS per wave of k_skinny_dot: its A tile(s) (PS of them with
// DMA), or the accumulators for the final sum
__host__ __device__ constexpr int dot_wave_floats(int rt, int ct) {
  return (dot_dma(rt, ct) ? dot_stages(rt, ct) : 1) * dot_tile_floats(rt) > rt * ct * 256
             ? (dot_dma(rt, ct) ? dot_stages(rt, ct) : 1) * dot_tile_floats(rt)
             : rt * ct * 256;
}
constexpr size_t kDotMaxLds = 163840;

// s_waitcnt immediates (gfx9 encoding: vmcnt [3:0] and [15:14], expcnt
// [6:4], lgkmcnt [11:8]); the other counters at their maximum (no wait)
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | 0x0F70; }
constexpr int kWaitLgkm0 = 0xC07F;

// A tile address (floats) of row x, 16-byte chunk kc
__device__ __forceinline__ int dot_tile_off(int x, int kc) {
  return x * kDotStageK + 4 * (kc ^ (x & 15));
}

// Block coordinates of the skinny kernels' (row block x K group, column
// block) grid.  Workgroups are dispatched in linear order L = y * gridDim.x
// + x and dealt to the 8 XCDs round robin (XCD L % 8, each with its own
// L2).  With several column blocks, each of which streams the same rows of
// A, L is read so that the column blocks of one (row block, K group) land
// on one XCD, dispatched 8 workgroups apart: the second reads of A come
// from that XCD's L2 instead of HBM.  (A bijection whenever gridDim.x is a
// multiple of 8; other grids keep the plain order.)
__device__ __forceinline__ void xcd_block(int64_t *bx, int *by) {
  const int64_t nx = gridDim.x;
  const int ny = (int)gridDim.y;
  if (ny == 1 || (nx & 7)) {
    *bx = blockIdx.x;
    *by = (int)blockIdx.y;
    return;
  }
  const int64_t L = (int64_t)blockIdx.y * nx + blockIdx.x;
  const int64_t j = L >> 3;
  *by = (int)(j % ny);
  *bx = (j / ny) * 8 + (L & 7);
}

__device__ __forceinline__ void dot_store(const DotParams &p, int64_t row,
                                          int64_t col, int kgi, float v) {
  if (row < p.M && col < p.N) {
    if (p.kgroups == 1) {
      float *cp = p.C + row * p.sCm + col * p.sCn;
      float o = p.alpha * v;
      if (p.read_c) o = p.alpha * v + p.beta * (*cp);
      *cp = o;
    } else {
      p.P[((int64_t)kgi * p.M + row) * p.N + col] = v;
    }
  }
}

// One wave: rows [row0, row0 + 16 RT) x columns [ct0*16, ct0*16 + 16 CT)
// over k in [kb, ke) (kb a multiple of the 64-k stage), accumulated into
// acc.  T is the wave's private LDS: one A tile (16 RT x 64 floats) when the
// stages go through registers, PS tiles when DMA.
//
// DMA: every stage of A goes from memory straight into its LDS tile
// (buffer_load_dwordx4 ... lds writes lane l's 16 bytes at the tile base +
// 1 KB q + 16 l), so no registers hold A in flight.  Lane (g, r) of
// instruction q loads row 4q + g, chunk r ^ ((4q + g) % 16): the same 4 rows
// x 256 bytes per instruction, landing in the swizzled tile layout.  The
// compiler does not order these LDS writes against the tile's reads, so the
// waits are explicit: before a stage's reads, vmcnt(NQ) (only the NQ loads
// of the stage issued last may still be in flight: the loads of the stage
// being read and its B came earlier), and before a tile is refilled,
// lgkmcnt(0) (its reads are done).
template <int RT, int CT, bool BVEC, bool DMA, int APOL = kBufNT>
__device__ __forceinline__ void dot_tile_lds(const float *__restrict__ A, int64_t sAm,
                                             const float *__restrict__ B, int64_t sBk,
                                             int64_t sBn, int64_t M, int64_t N,
                                             int64_t row0, int ct0, int64_t kb,
                                             int64_t ke, float *T, f4 (&acc)[RT][CT]) {
  constexpr int PS = dot_stages(RT, CT);
  constexpr int NQ = 4 * RT;  // A load instructions per stage
  constexpr int TF = dot_tile_floats(RT);
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
  const int64_t nst = ke > kb ? (ke - kb) / kDotStageK : 0;

  // A and B are read with buffer loads: a wave-uniform descriptor and
  // stage offset, plus a 32-bit byte offset per lane that carries every
  // lane-dependent part of the address.  A lane offset of kOOB is past any
  // descriptor's range, so the load returns zero without touching memory:
  // that is how rows past M, columns past N and k past the chunk end are
  // read.  Stages past the chunk load through a descriptor of size 0 (also
  // zero, no traffic), so the pipeline needs no guarded drain.  The host
  // guarantees every in-range offset of a wave is below 2^31.
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(A + row0 * sAm + kb, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rbs =
      make_rsrc(B + kb * sBk + (int64_t)ct0 * 16 * sBn, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rnull = make_rsrc(A, 0);
  uint32_t aoff[NQ];  // row row0 + 4q + g, k = kb + 4r (DMA: 4 (r ^ row % 16))
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int x = 4 * q + g;
    const int kc = DMA ? (r ^ (x & 15)) : r;
    aoff[q] = row0 + x < M ? (uint32_t)((x * sAm + 4 * kc) * 4) : kOOB;
  }
  uint32_t boff[CT];  // column ct0*16 + 16c + r, k = kb + 4g
#pragma unroll
  for (int c = 0; c < CT; ++c)
    boff[c] = (int64_t)(ct0 + c) * 16 + r < N
                  ? (uint32_t)((4 * g * sBk + (16 * c + r) * sBn) * 4) : kOOB;

#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = f4{0.f, 0.f, 0.f, 0.f};
  f4 a[DMA ? 1 : PS][NQ];
  f4 b[CT][2][2];  // one stage of B: [column tile][k-step u][half v]
  auto loadA = [&](int j, int64_t s) {
    const __amdgpu_buffer_rsrc_t rs = s < nst ? ra : rnull;
    const int32_t so = (int32_t)(s * kDotStageK * 4);
    if constexpr (DMA) {
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        buf_load_lds16<APOL>(rs, T + j * TF + q * 256, aoff[q], so);
    } else {
#pragma unroll
      for (int q = 0; q < NQ; ++q) a[j][q] = buf_load4<APOL>(rs, aoff[q], so);
    }
  };
  // k-step u (32 k) of stage s's B
  auto loadBu = [&](int u, int64_t s) {
    const __amdgpu_buffer_rsrc_t rs = s < nst ? rbs : rnull;
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int v = 0; v < 2; ++v) {
        const int64_t ko = s * kDotStageK + 32 * u + 16 * v;
        f4 x;
        if (BVEC) {
          x = buf_load4<0>(rs, boff[c], (int32_t)(ko * 4));
        } else {
#pragma unroll
          for (int jj = 0; jj < 4; ++jj)
            x[jj] = buf_load1(rs, boff[c], (int32_t)((ko + jj) * sBk * 4));
        }
        b[c][u][v] = x;
      }
  };
  // wave-local LDS ordering (the tile is private to the wave)
  auto wave_sync = [&]() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto put = [&](int j, float *Tj) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      *reinterpret_cast<f4 *>(Tj + dot_tile_off(4 * q + g, r)) = a[j][q];
    wave_sync();
  };
  // MFMAs of the stage in tile Tj and b; each k-step's B registers are
  // refilled with stage s_next's as soon as its MFMAs are issued
  auto compute = [&](const float *Tj, int64_t s_next) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      f4 x[RT][2];
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int v = 0; v < 2; ++v)
          x[t][v] = *reinterpret_cast<const f4 *>(Tj + dot_tile_off(16 * t + r, 8 * u + 4 * v + g));
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
          for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c)
              acc[t][c] = mfma4(x[t][v][jj], b[c][u][v][jj], acc[t][c]);
      loadBu(u, s_next);
    }
    // the tile may be overwritten only after these reads
    wave_sync();
  };

  if constexpr (DMA) {
    // Stage t lands in tile t % PS, PS stages ahead; B one stage ahead.
#pragma unroll
    for (int j = 0; j < PS; ++j) loadA(j, j);
    loadBu(0, 0);
    loadBu(1, 0);
    __builtin_amdgcn_sched_barrier(0);
    for (int64_t s = 0; s < nst; s += PS) {
#pragma unroll
      for (int j = 0; j < PS; ++j) {
        if (s + j < nst) {
          __builtin_amdgcn_s_waitcnt(waitcnt_vm(NQ));
          __builtin_amdgcn_sched_barrier(0);
          wave_sync();
          compute(T + j * TF, s + j + 1);
        }
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        __builtin_amdgcn_sched_barrier(0);
        loadA(j, s + PS + j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // every tile write (the trailing zero stages included) has landed
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    __builtin_amdgcn_sched_barrier(0);
    wave_sync();
  } else {
    // Full stages, PS at a time; stage t's A uses register set t % PS and
    // is loaded PS stages ahead (after the previous stage's B, so that
    // waiting for B never waits for later A); B is loaded one stage ahead.
#pragma unroll
    for (int j = 0; j < PS; ++j) loadA(j, j);
    loadBu(0, 0);
    loadBu(1, 0);
    __builtin_amdgcn_sched_barrier(0);
    for (int64_t s = 0; s < nst; s += PS) {
#pragma unroll
      for (int j = 0; j < PS; ++j) {
        const bool live = s + j < nst;
        if (live) {
          put(j, T);
          compute(T, s + j + 1);
        }
        loadA(j, s + PS + j);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }

  if (kb + nst * kDotStageK < ke) {
    // partial last stage, through registers: elements with k >= ke get the
    // out-of-range offset
    const int64_t k0 = kb + nst * kDotStageK;
    f4 ap[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const uint32_t base = row0 + 4 * q + g < M ? (uint32_t)(((4 * q + g) * sAm + 4 * r) * 4) : kOOB;
      f4 x;
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const int64_t k = k0 + 4 * r + jj;
        const uint32_t vo = (base == kOOB || k >= ke) ? kOOB
                            : base + (uint32_t)((nst * kDotStageK + jj) * 4);
        x[jj] = buf_load1(ra, vo, 0);
      }
      ap[q] = x;
    }
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
          f4 x;
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            const int64_t ko = nst * kDotStageK + 32 * u + 16 * v + jj;
            const uint32_t vo = (boff[c] == kOOB || kb + ko + 4 * g >= ke) ? kOOB
                                : boff[c] + (uint32_t)(ko * sBk * 4);
            x[jj] = buf_load1(rbs, vo, 0);
          }
          b[c][u][v] = x;
        }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      *reinterpret_cast<f4 *>(T + dot_tile_off(4 * q + g, r)) = ap[q];
    wave_sync();
    compute(T, nst + 1);  // nothing left to load: the refills read zeros, no traffic
  }
}

template <int RT, int CT, bool BVEC, int APOL = kBufNT>
__global__ __launch_bounds__(64 * dot_max_waves(RT, CT)) void k_skinny_dot(DotParams p) {
  constexpr int WF = dot_wave_floats(RT, CT);
  extern __shared__ f4 dot_lds[];
  // the wave index is uniform across the wave: say so, so that everything
  // derived from it (K range, descriptors) lives in scalar registers
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kw = blockDim.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  int64_t bx;
  int by;
  xcd_block(&bx, &by);
  const int64_t rb = bx / p.kgroups;
  const int kgi = (int)(bx - rb * p.kgroups);
  const int64_t row0 = rb * 16 * RT;
  const int ct0 = by * CT;
  const int64_t kb = ((int64_t)kgi * kw + w) * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;
  f4 acc[RT][CT];
  dot_tile_lds<RT, CT, BVEC, dot_dma(RT, CT), APOL>(p.A, p.sAm, p.B, p.sBk, p.sBn, p.M, p.N,
                                                   row0, ct0, kb, ke,
                                                   reinterpret_cast<float *>(dot_lds) + w * WF, acc);

  if (kw == 1) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          dot_store(p, row0 + 16 * t + 4 * g + i, (int64_t)(ct0 + c) * 16 + r, kgi, acc[t][c][i]);
    return;
  }
  // sum over the waves of the workgroup, in wave order (the tiles are done)
  __syncthreads();
  constexpr int NT = RT * CT;
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) dot_lds[(w * NT + t * CT + c) * 64 + lane] = acc[t][c];
  __syncthreads();
  for (int q = w; q < NT; q += kw) {
    f4 sum = dot_lds[q * 64 + lane];
    for (int ww = 1; ww < kw; ++ww) {
      const f4 x = dot_lds[(ww * NT + q) * 64 + lane];
      sum[0] += x[0]; sum[1] += x[1]; sum[2] += x[2]; sum[3] += x[3];
    }
    const int t = q / CT, c = q - t * CT;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      dot_store(p, row0 + 16 * t + 4 * g + i, (int64_t)(ct0 + c) * 16 + r, kgi, sum[i]);
  }
}

// ---------------------------------------------------------------------------
// Persistent grouped SGEMM: hundreds of independent skinny problems in one
// graph launch.  The host sorts the problems into tile classes by N (1, 2
// or 4 column tiles of 16 per pass), builds one descriptor per problem and,
// per class, a list of (problem, row block) tiles sorted by decreasing
// cost; one persistent kernel per class runs its list: wave w takes tiles
// w, w + W, w + 2W, ... (W = waves launched), a static longest-first
// schedule with no atomics or counters, so the launches can be captured in
// a graph and replayed.  Each tile runs the DOT tile body above (B read in
// place, whole K, no split) over its problem's column tiles, RT row tiles
// of 16 per wave (2, or 1 for more resident waves).  One kernel per class
// keeps each at its own register budget (a single kernel dispatching on
// the class would pay for the largest); the 4-column class stages A through
// LDS DMA, which is what lets its 2 x 4 tile fit 2 waves per SIMD without
// spilling.
// ---------------------------------------------------------------------------
// Stores a wave's RT x CT tile of one grouped problem, C = alpha acc
// (+ beta C).  The descriptor's fields are read into registers once: as far
// as the compiler knows the C stores may alias the descriptor array, so
// fields read through the reference inside the loop were reloaded from
// memory after every store, one dependent round trip per element.
template <int RT, int CT>
__device__ __forceinline__ void grouped_store(const GroupProblem &gp, int64_t row0, int ct0,
                                              const f4 (&acc)[RT][CT]) {
  float *const C = gp.C;
  const int64_t ldc = gp.ldc, M = gp.M, N = gp.N;
  const float alpha = gp.alpha, beta = gp.beta;
  const bool read_c = gp.read_c != 0;
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, g = lane >> 4;
#pragma unroll
  for (int tt = 0; tt < RT; ++tt)
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t row = row0 + 16 * tt + 4 * g + i;
        const int64_t col = (int64_t)(ct0 + c) * 16 + r;
        if (row < M && col < N) {
          float *cp = C + row * ldc + col;
          float v = alpha * acc[tt][c][i];
          if (read_c) v = alpha * acc[tt][c][i] + beta * (*cp);
          *cp = v;
        }
      }
}

template <int RT, int CT>
__global__ __launch_bounds__(64 * kGroupWaveWpb, 2) void k_grouped_dot(const GroupProblem *probs,
                                                     const GroupTile *tiles,
                                                     int64_t ntiles) {
  __shared__ f4 gtile[kGroupWaveWpb][(dot_dma(RT, CT) ? dot_stages(RT, CT) : 1) * dot_tile_floats(RT) / 4];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  const int64_t W = (int64_t)gridDim.x * (blockDim.x >> 6);
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + wv;
  float *T = reinterpret_cast<float *>(gtile[wv]);
  for (int64_t t = w; t < ntiles; t += W) {
    const GroupTile tile = tiles[t];
    if (tile.prob < 0) continue;  // a hole of the list's layout
    const GroupProblem &gp = probs[tile.prob];
    const int64_t row0 = tile.rb * 16 * RT;
    const int64_t nct = (gp.N + 15) / 16;
    for (int ct0 = 0; ct0 < nct; ct0 += CT) {
      f4 acc[RT][CT];
      dot_tile_lds<RT, CT, false, dot_dma(RT, CT), 0>(gp.A, gp.lda, gp.B, gp.ldb, 1, gp.M, gp.N,
                                                     row0, ct0, 0, gp.K, T, acc);
      grouped_store<RT, CT>(gp, row0, ct0, acc);
    }
  }
}

// ---------------------------------------------------------------------------
// AXPY form: A(m,k) = A[m + k*sAk] (contiguous along M).  A wave computes RT
// row tiles of 64 rows x CT column tiles over its own K chunk; the kw waves
// of a workgroup take consecutive chunks of the same tiles and sum their
// results through LDS at the end, in wave order, so a K split needs no
// second kernel unless it spans workgroups.
//
// K advances in steps of 16.  Lane (g, r) loads A(m0 + 4r .. m0 + 4r + 3, k)
// for k = k0 + 4g + u, u = 0..3 -- per load instruction 4 columns x 256
// contiguous bytes.  MFMA (u, j) takes element j of load u: its 16 "rows"
// are m0 + 4r + j (r = 0..15) and its 4 k are k0 + 4g + u (g = 0..3).  Its
// B operand at lane (g, r) of column tile c is B(k0 + 4g + u, 16c + r); a
// lane's 4 values (u = 0..3) are consecutive in k, so K-contiguous B (BVEC)
// is one 16-byte load per column tile and step, other layouts four scalar
// loads.  B is read in place with its own strides.  Result element i at
// lane (g, r) of MFMA (u, j), column tile c, is C(m0 + 4(4g + i) + j,
// 16c + r).
//
// Loads go through buffer descriptors rebuilt per step (base = the step's
// first column), so only one step's extent has to fit the 31-bit offsets.
// A lane offset of kOOB returns zero without touching memory: rows past M,
// columns past N and k past the chunk end.  Steps past the chunk load
// through a descriptor of size 0, so the pipeline needs no guarded drain.
// VEC (M % 4 == 0): a lane's 4 rows are all in range or all out, one 16-byte
// load; otherwise 4 scalar loads, each row checked on its own.
// ---------------------------------------------------------------------------
// register stages of loads in flight, and workgroup size limit, per shape
__host__ __device__ constexpr int axpy_stages(int rt, int ct) {
  return (rt == 1 && ct <= 2) ? 3 : 2;
}
__host__ __device__ constexpr int axpy_max_waves(int rt, int ct) {
  return rt * ct >= 8 ? 4 : 8;
}
constexpr int kAxpyRedGroups = 4;  // accumulator f4 groups summed per LDS pass

__device__ __forceinline__ void store_out(const SkinnyParams &p, int64_t row,
                                          int64_t col, int kgi, float v) {
  if (row < p.M && col < p.N) {
    if (p.kgroups == 1) {
      float *cp = p.C + row * p.sCm + col * p.sCn;
      float o = p.alpha * v;
      if (p.read_c) o = p.alpha * v + p.beta * (*cp);
      *cp = o;
    } else {
      p.P[((int64_t)kgi * p.M + row) * p.N + col] = v;
    }
  }
}

template <int RT, int CT, bool VEC, bool BVEC>
__global__ __launch_bounds__(64 * axpy_max_waves(RT, CT)) void k_skinny_axpy(SkinnyParams p) {
  constexpr int PS = axpy_stages(RT, CT);
  extern __shared__ f4 axpy_lds[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int kw = blockDim.x >> 6;
  const int r = lane & 15, g = lane >> 4;
  int64_t bx;
  int by;
  xcd_block(&bx, &by);
  const int64_t rb = bx / p.kgroups;
  const int kgi = (int)(bx - rb * p.kgroups);
  const int64_t m0 = rb * kAxpyRowsPerTile * RT;
  const int ct0 = by * CT;
  const int64_t kb = ((int64_t)kgi * kw + w) * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;
  const int64_t nfull = ke > kb ? (ke - kb) / kAxpyKStep : 0;

  // lane offsets (bytes) from a step's descriptor bases
  uint32_t aoff[RT][VEC ? 1 : 4];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int e = 0; e < (VEC ? 1 : 4); ++e) {
      const int64_t row = 64 * t + 4 * r + e;
      aoff[t][e] = m0 + row < p.M ? (uint32_t)((row + 4 * g * p.sAk) * 4) : kOOB;
    }
  uint32_t boff[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c)
    boff[c] = (int64_t)(ct0 + c) * 16 + r < p.N
                  ? (uint32_t)(((16 * c + r) * p.sBn + 4 * g * p.sBk) * 4) : kOOB;
  const float *Aw = p.A + kb * p.sAk + m0;
  const float *Bw = p.B + kb * p.sBk + (int64_t)ct0 * 16 * p.sBn;
  const __amdgpu_buffer_rsrc_t rnull = make_rsrc(p.A, 0);
  const int64_t astep = (int64_t)kAxpyKStep * p.sAk, bstep = (int64_t)kAxpyKStep * p.sBk;

  f4 acc[RT][4][CT];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][j][c] = f4{0.f, 0.f, 0.f, 0.f};

  f4 a[PS][RT][4], b[PS][CT];
  auto load = [&](int j, int64_t s) {
    const bool live = s < nfull;
    const __amdgpu_buffer_rsrc_t ra = live ? make_rsrc(Aw + s * astep, kRsrcMax) : rnull;
    const __amdgpu_buffer_rsrc_t rbs = live ? make_rsrc(Bw + s * bstep, kRsrcMax) : rnull;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t so = (int32_t)(u * p.sAk * 4);
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        if (VEC) {
          a[j][t][u] = buf_load4<kBufNT>(ra, aoff[t][0], so);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) a[j][t][u][e] = buf_load1<kBufNT>(ra, aoff[t][e], so);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      if (BVEC) {
        b[j][c] = buf_load4<0>(rbs, boff[c], 0);
      } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) b[j][c][u] = buf_load1(rbs, boff[c], (int32_t)(u * p.sBk * 4));
      }
    }
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

  // full steps, PS at a time: step t uses register set t % PS and is loaded
  // PS steps ahead, A and B of a step together (one in-order load counter)
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
  if (kb + nfull * kAxpyKStep < ke) {
    // partial last step: k >= ke reads as zero (per lane: its k is k0 + 4g + u)
    const int64_t k0 = kb + nfull * kAxpyKStep;
    const __amdgpu_buffer_rsrc_t ra = make_rsrc(Aw + nfull * astep, kRsrcMax);
    const __amdgpu_buffer_rsrc_t rbs = make_rsrc(Bw + nfull * bstep, kRsrcMax);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool kin = k0 + 4 * g + u < ke;
      const int32_t so = (int32_t)(u * p.sAk * 4);
#pragma unroll
      for (int t = 0; t < RT; ++t) {
        if (VEC) {
          a[0][t][u] = buf_load4<kBufNT>(ra, kin ? aoff[t][0] : kOOB, so);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            a[0][t][u][e] = buf_load1<kBufNT>(ra, kin ? aoff[t][e] : kOOB, so);
        }
      }
#pragma unroll
      for (int c = 0; c < CT; ++c)
        b[0][c][u] = buf_load1(rbs, kin ? boff[c] : kOOB, (int32_t)(u * p.sBk * 4));
    }
    comp(0);
  }

  if (kw == 1) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int c = 0; c < CT; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i)
            store_out(p, m0 + 64 * t + 4 * (4 * g + i) + j, (int64_t)(ct0 + c) * 16 + r, kgi,
                      acc[t][j][c][i]);
    return;
  }
  // Sum over the waves of the workgroup in wave order, kAxpyRedGroups
  // accumulator groups (group q = (t*4 + j)*CT + c) per pass: every wave
  // writes its groups, then wave w sums groups w, w + kw, ... of the pass.
  constexpr int NG = RT * 4 * CT;
#pragma unroll
  for (int pass = 0; pass < NG / kAxpyRedGroups; ++pass) {
    __syncthreads();  // the previous pass's reads are done
#pragma unroll
    for (int i = 0; i < kAxpyRedGroups; ++i) {
      const int q = pass * kAxpyRedGroups + i;
      axpy_lds[(w * kAxpyRedGroups + i) * 64 + lane] = acc[q / (4 * CT)][(q / CT) % 4][q % CT];
    }
    __syncthreads();
    for (int i = w; i < kAxpyRedGroups; i += kw) {
      f4 sum = axpy_lds[i * 64 + lane];
      for (int ww = 1; ww < kw; ++ww) {
        const f4 x = axpy_lds[(ww * kAxpyRedGroups + i) * 64 + lane];
        sum[0] += x[0]; sum[1] += x[1]; sum[2] += x[2]; sum[3] += x[3];
      }
      const int q = pass * kAxpyRedGroups + i;
      const int t = q / (4 * CT), j = (q / CT) % 4, c = q % CT;
#pragma unroll
      for (int ii = 0; ii < 4; ++ii)
        store_out(p, m0 + 64 * t + 4 * (4 * g + ii) + j, (int64_t)(ct0 + c) * 16 + r, kgi,
                  sum[ii]);
    }
  }
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


// ---------------------------------------------------------------------------
// DOT form with B shared through LDS: k_skinny_dotb.
//
// The 4 waves of a workgroup take 4 consecutive row tiles (16 RT rows
// each) of one column block and one K chunk, so a single copy of every
// 64-k stage of B serves all four.  The workgroup loads the stage
// cooperatively (CT 16-byte loads per thread) and stores it with column n's
// 64 k values together, 16-byte chunk kq at n*64 + 4*(kq ^ (n % 16)), from
// where each wave reads its MFMA fragments with conflict-free 16-byte LDS
// reads.  A goes through each wave's private LDS tile as in k_skinny_dot.
// Registers hold 2 stages of A and of the thread's share of B (CT loads
// instead of the 16 CT a wave needs for its own copy), which keeps wide
// column blocks at two or more waves per SIMD, and B's L2 traffic falls by
// the 4 waves.  One workgroup barrier per stage; a K split spans workgroups
// (partials summed in a fixed order by k_reduce).
// ---------------------------------------------------------------------------
constexpr int kDotbWaves = 4;
// LDS bytes of k_skinny_dotb<rt, ct>: 4 A tiles and 2 B stages
constexpr size_t dotb_lds(int rt, int ct) {
  return (size_t)(kDotbWaves * rt + 2 * ct) * 16 * kDotStageK * sizeof(float);
}

__device__ __forceinline__ int dotb_off(int n, int kq) { return n * kDotStageK + 4 * (kq ^ (n & 15)); }

// One workgroup of 4 waves: wave w computes rows [row0, row0 + 16 RT)
// (row0 = block row + 16 RT w) x columns [16 ct0, 16 ct0 + 16 CT) over k in
// [kb, ke) into acc, the block's B stages shared through LDS.  Every wave
// of the workgroup must call it with the same ct0, kb and ke (it holds
// workgroup barriers); lds holds dotb_lds(RT, CT) bytes.  On return the
// LDS may still be read by other waves: barrier before reusing it.
template <int RT, int CT, bool BVEC, int APOL = kBufNT>
__device__ __forceinline__ void dotb_block(const float *__restrict__ A, int64_t sAm,
                                           const float *__restrict__ B, int64_t sBk,
                                           int64_t sBn, int64_t M, int64_t N, int64_t row0,
                                           int ct0, int64_t kb, int64_t ke, float *lds,
                                           f4 (&acc)[RT][CT]) {
  constexpr int PS = 2;
  constexpr int NQ = 4 * RT;        // A loads per stage per lane
  constexpr int TW = RT * 16 * kDotStageK;  // floats of a wave's A tile
  constexpr int BW = CT * 16 * kDotStageK;  // floats of one B stage
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int r = lane & 15, g = lane >> 4;
  float *T = lds + w * TW;
  float *Bs = lds + kDotbWaves * TW;  // [2][16 CT][64]
  const int64_t n0 = (int64_t)ct0 * 16;
  const int64_t nst = ke > kb ? (ke - kb) / kDotStageK : 0;
  const bool tail = kb + nst * kDotStageK < ke;

  // A: as in dot_tile_lds (rows past M read zero)
  const __amdgpu_buffer_rsrc_t ra = make_rsrc(A + row0 * sAm + kb, kRsrcMax);
  const __amdgpu_buffer_rsrc_t rnull = make_rsrc(A, 0);
  uint32_t aoff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
    aoff[q] = row0 + 4 * q + g < M ? (uint32_t)(((4 * q + g) * sAm + 4 * r) * 4) : kOOB;
  // B: thread tid's CT 16-byte pieces of a stage.  K-contiguous B: piece
  // f = tid + 256 i is column n = f / 16, k-quad f % 16.  N-contiguous B:
  // k = f / (4 CT), columns 4 (f % 4CT) .. +3, stored element by element.
  const __amdgpu_buffer_rsrc_t rb = make_rsrc(BVEC ? B + kb + n0 * sBn : B + kb * sBk + n0, kRsrcMax);
  uint32_t boff[CT];
  int bn[CT], bk[CT];
  bool bwhole[CT];
#pragma unroll
  for (int i = 0; i < CT; ++i) {
    const int f = tid + 256 * i;
    if (BVEC) {
      bn[i] = f >> 4;
      bk[i] = 4 * (f & 15);
      bwhole[i] = true;
      boff[i] = n0 + bn[i] < N ? (uint32_t)((bn[i] * sBn + bk[i]) * 4) : kOOB;
    } else {
      bk[i] = f / (4 * CT);
      bn[i] = 4 * (f % (4 * CT));
      bwhole[i] = n0 + bn[i] + 3 < N;
      boff[i] = n0 + bn[i] < N ? (uint32_t)((bk[i] * sBk + bn[i]) * 4) : kOOB;
    }
  }

#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[t][c] = f4{0.f, 0.f, 0.f, 0.f};
  f4 a[PS][NQ];
  f4 b[PS][CT];
  auto loadA = [&](int j, int64_t s) {
    const __amdgpu_buffer_rsrc_t rs = s < nst ? ra : rnull;
    const int32_t so = (int32_t)(s * kDotStageK * 4);
#pragma unroll
    for (int q = 0; q < NQ; ++q) a[j][q] = buf_load4<APOL>(rs, aoff[q], so);
  };
  auto loadB = [&](int j, int64_t s) {
    const __amdgpu_buffer_rsrc_t rs = s < nst ? rb : rnull;
    const int32_t so = (int32_t)(s * kDotStageK * (BVEC ? 1 : sBk) * 4);
#pragma unroll
    for (int i = 0; i < CT; ++i) {
      if (bwhole[i]) {
        b[j][i] = buf_load4<0>(rs, boff[i], so);
      } else {
        f4 x;
#pragma unroll
        for (int e = 0; e < 4; ++e)
          x[e] = buf_load1(rs, n0 + bn[i] + e < N ? boff[i] + 4 * e : kOOB, so);
        b[j][i] = x;
      }
    }
  };
  // stage j's registers into LDS (A tile of the wave, B buffer `buf`), then
  // a workgroup barrier
  auto put = [&](int j, int buf) {
#pragma unroll
    for (int q = 0; q < NQ; ++q)
      *reinterpret_cast<f4 *>(T + dot_tile_off(4 * q + g, r)) = a[j][q];
    float *B2 = Bs + buf * BW;
#pragma unroll
    for (int i = 0; i < CT; ++i) {
      if (BVEC) {
        *reinterpret_cast<f4 *>(B2 + dotb_off(bn[i], bk[i] >> 2)) = b[j][i];
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
          B2[dotb_off(bn[i] + e, bk[i] >> 2) + (bk[i] & 3)] = b[j][i][e];
      }
    }
    __syncthreads();
  };
  auto compute = [&](int buf) {
    const float *B2 = Bs + buf * BW;
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      f4 x[RT][2], y[CT][2];
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int v = 0; v < 2; ++v)
          x[t][v] = *reinterpret_cast<const f4 *>(T + dot_tile_off(16 * t + r, 8 * u + 4 * v + g));
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int v = 0; v < 2; ++v)
          y[c][v] = *reinterpret_cast<const f4 *>(B2 + dotb_off(16 * c + r, 8 * u + 4 * v + g));
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
          for (int t = 0; t < RT; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c)
              acc[t][c] = mfma4(x[t][v][jj], y[c][v][jj], acc[t][c]);
    }
    // the wave's next put() may overwrite its A tile only after these reads
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

#pragma unroll
  for (int j = 0; j < PS; ++j) {
    loadA(j, j);
    loadB(j, j);
    __builtin_amdgcn_sched_barrier(0);
  }
  // nst and the tail are the same for the whole workgroup: the barriers in
  // put() are reached by every wave
  for (int64_t s = 0; s < nst; s += PS) {
#pragma unroll
    for (int j = 0; j < PS; ++j) {
      const bool live = s + j < nst;
      if (live) put(j, (int)((s + j) & 1));
      loadA(j, s + PS + j);
      loadB(j, s + PS + j);
      if (live) compute((int)((s + j) & 1));
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (tail) {
    // partial last stage, element by element past ke
    const int64_t k0 = nst * kDotStageK;  // relative to kb
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      f4 x;
#pragma unroll
      for (int jj = 0; jj < 4; ++jj) {
        const uint32_t vo = (aoff[q] == kOOB || kb + k0 + 4 * r + jj >= ke) ? kOOB
                            : aoff[q] + (uint32_t)((k0 + jj) * 4);
        x[jj] = buf_load1(ra, vo, 0);
      }
      a[0][q] = x;
    }
#pragma unroll
    for (int i = 0; i < CT; ++i) {
      f4 x;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t k = BVEC ? k0 + bk[i] + e : k0 + bk[i];
        const int64_t n = n0 + bn[i] + (BVEC ? 0 : e);
        uint32_t vo = kOOB;
        if (n < N && kb + k < ke)
          vo = BVEC ? (uint32_t)((bn[i] * sBn + k) * 4) : (uint32_t)((k * sBk + bn[i] + e) * 4);
        x[e] = buf_load1(rb, vo, 0);
      }
      b[0][i] = x;
    }
    put(0, (int)(nst & 1));
    compute((int)(nst & 1));
  }

}

template <int RT, int CT, bool BVEC>
__global__ __launch_bounds__(64 * kDotbWaves) void k_skinny_dotb(DotParams p) {
  extern __shared__ f4 dotb_smem[];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  int64_t bx;
  int by;
  xcd_block(&bx, &by);
  const int64_t rbk = bx / p.kgroups;
  const int kgi = (int)(bx - rbk * p.kgroups);
  const int64_t row0 = (rbk * kDotbWaves + w) * 16 * RT;
  const int ct0 = by * CT;
  const int64_t kb = (int64_t)kgi * p.kchunk;
  int64_t ke = kb + p.kchunk;
  if (ke > p.K) ke = p.K;
  f4 acc[RT][CT];
  dotb_block<RT, CT, BVEC>(p.A, p.sAm, p.B, p.sBk, p.sBn, p.M, p.N, row0, ct0, kb, ke,
                           reinterpret_cast<float *>(dotb_smem), acc);
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        dot_store(p, row0 + 16 * t + 4 * g + i, (int64_t)(ct0 + c) * 16 + r, kgi, acc[t][c][i]);
}


// Grouped SGEMM with B shared through LDS: the work list holds blocks of 4
// consecutive row tiles (16 RT rows each) of one problem; a workgroup takes
// blocks b, b + G, ... (G workgroups launched, a static longest-first
// schedule as in k_grouped_dot) and runs dotb_block on each, its 4 waves one
// row tile each, so every 64-k stage of the problem's B is read from L2 once
// per block instead of once per row tile.  Row-major problems only (B is
// N-contiguous).
template <int RT, int CT>
__global__ __launch_bounds__(64 * kDotbWaves, 2) void k_grouped_dotb(const GroupProblem *probs,
                                                                  const GroupTile *tiles,
                                                                  int64_t ntiles) {
  extern __shared__ f4 gdotb_smem[];
  float *lds = reinterpret_cast<float *>(gdotb_smem);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 15, g = lane >> 4;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const GroupTile tile = tiles[t];
    if (tile.prob < 0) continue;  // a hole of the list's layout
    const GroupProblem &gp = probs[tile.prob];
    const int64_t row0 = (tile.rb * kDotbWaves + w) * 16 * RT;
    const int64_t nct = (gp.N + 15) / 16;
    for (int ct0 = 0; ct0 < nct; ct0 += CT) {
      f4 acc[RT][CT];
      dotb_block<RT, CT, false, 0>(gp.A, gp.lda, gp.B, gp.ldb, 1, gp.M, gp.N, row0, ct0, 0,
                                   gp.K, lds, acc);
      grouped_store<RT, CT>(gp, row0, ct0, acc);
      // the next block's first stage overwrites the LDS other waves may
      // still be reading
      __syncthreads();
    }
  }
}

template <int RT, int CT, bool BV, int APOL>
hipError_t run_dot(int kw, const DotParams &p, hipStream_t st) {
  const int64_t nrb = (p.M + 16 * RT - 1) / (16 * RT);
  const int64_t nct = (p.N + 15) / 16;
  const int64_t ncb = (nct + CT - 1) / CT;
  const int64_t nblk = nrb * p.kgroups;
  if (nblk > 0x7fffffff || ncb > 65535) return hipErrorInvalidConfiguration;
  const size_t lds = (size_t)kw * dot_wave_floats(RT, CT) * sizeof(float);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k_skinny_dot<RT, CT, BV, APOL>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_skinny_dot<RT, CT, BV, APOL>), dim3((unsigned)nblk, (unsigned)ncb),
                     dim3(64 * kw), lds, st, p);
  return hipGetLastError();
}

typedef hipError_t (*DotFn)(int, const DotParams &, hipStream_t);

// cached: A read with ordinary (cached) loads instead of non-temporal ones
template <int RT, int CT>
DotFn dot_fn(bool bv, bool cached) {
  return cached ? (bv ? run_dot<RT, CT, true, 0> : run_dot<RT, CT, false, 0>)
                : (bv ? run_dot<RT, CT, true, kBufNT> : run_dot<RT, CT, false, kBufNT>);
}

// instantiated (rt, ct): ct <= 4, rt * ct <= 8
DotFn find_dot(int rt, int ct, bool bv, bool cached = false) {
  switch (rt * 16 + ct) {
    case 17: return dot_fn<1, 1>(bv, cached);
    case 18: return dot_fn<1, 2>(bv, cached);
    case 20: return dot_fn<1, 4>(bv, cached);
    case 33: return dot_fn<2, 1>(bv, cached);
    case 34: return dot_fn<2, 2>(bv, cached);
    case 36: return dot_fn<2, 4>(bv, cached);
    case 65: return dot_fn<4, 1>(bv, cached);
    case 66: return dot_fn<4, 2>(bv, cached);
    default: return nullptr;
  }
}




template <int RT, int CT, bool BV>
hipError_t run_dotb(const DotParams &p, hipStream_t st) {
  const int64_t nrb = (p.M + kDotbWaves * 16 * RT - 1) / (kDotbWaves * 16 * RT);
  const int64_t nct = (p.N + 15) / 16;
  const int64_t ncb = (nct + CT - 1) / CT;
  const int64_t nblk = nrb * p.kgroups;
  if (nblk > 0x7fffffff || ncb > 65535) return hipErrorInvalidConfiguration;
  const size_t lds = dotb_lds(RT, CT);
  hipLaunchKernelGGL((k_skinny_dotb<RT, CT, BV>), dim3((unsigned)nblk, (unsigned)ncb),
                     dim3(64 * kDotbWaves), lds, st, p);
  return hipGetLastError();
}

typedef hipError_t (*DotbFn)(const DotParams &, hipStream_t);

template <int RT, int CT>
DotbFn dotb_fn(bool bv) {
  return bv ? run_dotb<RT, CT, true> : run_dotb<RT, CT, false>;
}

// instantiated (rt, ct): rt in {1, 2}, ct in {1, 2, 4} (4 row tiles of 4
// take more than 64 KB of LDS)
DotbFn find_dotb(int rt, int ct, bool bv) {
  switch (rt * 16 + ct) {
    case 17: return dotb_fn<1, 1>(bv);
    case 18: return dotb_fn<1, 2>(bv);
    case 20: return dotb_fn<1, 4>(bv);
    case 33: return dotb_fn<2, 1>(bv);
    case 34: return dotb_fn<2, 2>(bv);
    case 36: return dotb_fn<2, 4>(bv);
    default: return nullptr;
  }
}

// Workgroups (256 threads, dynamic LDS `lds`) of kernel `fn` the current
// device holds at once, from the occupancy calculator; 0 if unknown.
int resident_blocks(const void *fn, size_t lds, int threads = 256) {
  int nb = 0, dev = 0;
  hipDeviceProp_t prop;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, threads, lds) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return std::max(1, nb) * prop.multiProcessorCount;
}

template <int RT, int CT, bool VEC, bool BVEC>
hipError_t run_axpy(int kw, const SkinnyParams &p, hipStream_t st) {
  const int64_t rows = (int64_t)kAxpyRowsPerTile * RT;
  const int64_t nblk = ((p.M + rows - 1) / rows) * p.kgroups;
  const int64_t ncb = p.nct_pad / CT;
  if (nblk > 0x7fffffff || ncb > 65535) return hipErrorInvalidConfiguration;
  const size_t lds = kw > 1 ? (size_t)kw * kAxpyRedGroups * 64 * sizeof(f4) : 0;
  hipLaunchKernelGGL((k_skinny_axpy<RT, CT, VEC, BVEC>), dim3((unsigned)nblk, (unsigned)ncb),
                     dim3(64 * kw), lds, st, p);
  return hipGetLastError();
}

typedef hipError_t (*RunFn)(int, const SkinnyParams &, hipStream_t);

template <int RT, int CT>
constexpr RunFn axpy_fn(bool vec, bool bvec) {
  return vec ? (bvec ? run_axpy<RT, CT, true, true> : run_axpy<RT, CT, true, false>)
             : (bvec ? run_axpy<RT, CT, false, true> : run_axpy<RT, CT, false, false>);
}

RunFn find_axpy(int rt, int ct, bool vec, bool bvec) {
  // 2 x 64 rows x 128 columns would spill: no (2, 8) tile
  switch (rt * 16 + ct) {
    case 17: return axpy_fn<1, 1>(vec, bvec);
    case 18: return axpy_fn<1, 2>(vec, bvec);
    case 20: return axpy_fn<1, 4>(vec, bvec);
    case 24: return axpy_fn<1, 8>(vec, bvec);
    case 33: return axpy_fn<2, 1>(vec, bvec);
    case 34: return axpy_fn<2, 2>(vec, bvec);
    case 36: return axpy_fn<2, 4>(vec, bvec);
    default: return nullptr;
  }
}

}  // namespace

bool skinny_has_variant(int form, int rt, int ct) {
  if (form == kFormDot) return find_dot(rt, ct, true) != nullptr;
  return find_axpy(rt, ct, true, true) != nullptr;
}

size_t skinny_dot_lds_bytes(int rt, int ct, int kw) {
  return (size_t)kw * dot_wave_floats(rt, ct) * sizeof(float);
}

int skinny_dot_resident_waves(int rt, int ct) {
  // waves of k_skinny_dot<rt, ct> (4-wave workgroups, B read with scalar
  // loads) that the current device holds at once; cached
  static std::mutex mu;
  static int cache[5][5];
  if (rt < 1 || rt > 4 || ct < 1 || ct > 4 || !find_dot(rt, ct, false)) return 0;
  std::lock_guard<std::mutex> lk(mu);
  if (cache[rt][ct]) return cache[rt][ct];
  const void *fn = nullptr;
  switch (rt * 16 + ct) {
    case 17: fn = (const void *)k_skinny_dot<1, 1, false>; break;
    case 18: fn = (const void *)k_skinny_dot<1, 2, false>; break;
    case 20: fn = (const void *)k_skinny_dot<1, 4, false>; break;
    case 33: fn = (const void *)k_skinny_dot<2, 1, false>; break;
    case 34: fn = (const void *)k_skinny_dot<2, 2, false>; break;
    case 36: fn = (const void *)k_skinny_dot<2, 4, false>; break;
    case 65: fn = (const void *)k_skinny_dot<4, 1, false>; break;
    case 66: fn = (const void *)k_skinny_dot<4, 2, false>; break;
    default: return 0;
  }
  cache[rt][ct] = 4 * resident_blocks(fn, skinny_dot_lds_bytes(rt, ct, 4));
  return cache[rt][ct];
}

bool skinny_dot_runnable(int rt, int ct, int kw) {
  if (!find_dot(rt, ct, true)) return false;
  return kw >= 1 && kw <= dot_max_waves(rt, ct) && skinny_dot_lds_bytes(rt, ct, kw) <= kDotMaxLds;
}

hipError_t launch_skinny_dot(int rt, int ct, bool bvec, int kw, const DotParams &p,
                             hipStream_t st) {
  if (!skinny_dot_runnable(rt, ct, kw) || p.kgroups < 1 || p.kchunk % kDotStageK != 0)
    return hipErrorInvalidValue;
  const int64_t ncb = ((p.N + 15) / 16 + ct - 1) / ct;
  DotFn fn = find_dot(rt, ct, bvec, a_loads_cached(p.A, p.sAm, ncb));
  return fn(kw, p, st);
}
bool skinny_dotb_runnable(int rt, int ct) {
  return find_dotb(rt, ct, true) != nullptr && dotb_lds(rt, ct) <= 65536;
}

hipError_t launch_skinny_dotb(int rt, int ct, bool bvec, const DotParams &p, hipStream_t st) {
  if (!skinny_dotb_runnable(rt, ct) || p.kgroups < 1 || p.kchunk % kDotStageK != 0)
    return hipErrorInvalidValue;
  return find_dotb(rt, ct, bvec)(p, st);
}

bool skinny_axpy_runnable(int rt, int ct, int kw) {
  return find_axpy(rt, ct, true, true) != nullptr && kw >= 1 && kw <= axpy_max_waves(rt, ct) &&
         (kw & (kw - 1)) == 0;
}

hipError_t launch_skinny_axpy(int rt, int ct, bool vec, bool bvec, int kw,
                              const SkinnyParams &p, hipStream_t st) {
  if (!skinny_axpy_runnable(rt, ct, kw) || p.nct_pad % ct != 0 || p.kgroups < 1 ||
      p.kchunk % kAxpyKStep != 0)
    return hipErrorInvalidValue;
  return find_axpy(rt, ct, vec, bvec)(kw, p, st);
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

namespace {
// Kernel of (class, kind): index kind * kGroupClasses + class.  The
// shared-B kind runs 1-row-tile tiles in the 3- and 4-column classes (2 x 3
// and 2 x 4 need more than the 256 registers of 2 waves per SIMD without
// the per-wave kernel's LDS staging of A).
constexpr int kGroupRt[kGroupKinds][kGroupClasses] = {{2, 2, 2, 2}, {1, 1, 1, 1}, {2, 2, 1, 1}};
constexpr int kGroupCt[kGroupClasses] = {1, 2, 3, 4};

template <int RT, int CT>
hipError_t launch_gwave(dim3 grid, const GroupProblem *probs, const GroupTile *tiles,
                        int64_t ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_grouped_dot<RT, CT>), grid, dim3(64 * kGroupWaveWpb), 0, st, probs, tiles,
                     ntiles);
  return hipGetLastError();
}
template <int RT, int CT>
hipError_t launch_gshared(dim3 grid, const GroupProblem *probs, const GroupTile *tiles,
                          int64_t ntiles, hipStream_t st) {
  hipLaunchKernelGGL((k_grouped_dotb<RT, CT>), grid, dim3(64 * kDotbWaves), dotb_lds(RT, CT), st,
                     probs, tiles, ntiles);
  return hipGetLastError();
}
typedef hipError_t (*GroupLaunch)(dim3, const GroupProblem *, const GroupTile *, int64_t,
                                  hipStream_t);
struct GroupKernel {
  const void *fn;
  GroupLaunch launch;
};
const GroupKernel kGroupKernels[kGroupKinds * kGroupClasses] = {
    {(const void *)k_grouped_dot<2, 1>, launch_gwave<2, 1>},
    {(const void *)k_grouped_dot<2, 2>, launch_gwave<2, 2>},
    {(const void *)k_grouped_dot<2, 3>, launch_gwave<2, 3>},
    {(const void *)k_grouped_dot<2, 4>, launch_gwave<2, 4>},
    {(const void *)k_grouped_dot<1, 1>, launch_gwave<1, 1>},
    {(const void *)k_grouped_dot<1, 2>, launch_gwave<1, 2>},
    {(const void *)k_grouped_dot<1, 3>, launch_gwave<1, 3>},
    {(const void *)k_grouped_dot<1, 4>, launch_gwave<1, 4>},
    {(const void *)k_grouped_dotb<2, 1>, launch_gshared<2, 1>},
    {(const void *)k_grouped_dotb<2, 2>, launch_gshared<2, 2>},
    {(const void *)k_grouped_dotb<1, 3>, launch_gshared<1, 3>},
    {(const void *)k_grouped_dotb<1, 4>, launch_gshared<1, 4>},
};
size_t grouped_lds(int cls, int kind) {
  return kind == kGroupShared ? dotb_lds(kGroupRt[kind][cls], kGroupCt[cls]) : 0;
}
}  // namespace

void grouped_kind_shape(int cls, int kind, int *rt, int *ct) {
  *rt = kGroupRt[kind][cls];
  *ct = kGroupCt[cls];
}

int grouped_unit_rows(int cls, int kind) {
  return 16 * kGroupRt[kind][cls] * (kind == kGroupShared ? kDotbWaves : 1);
}

int grouped_class_blocks(int cls, int kind) {
  static std::mutex mu;
  static int cache[kGroupKinds * kGroupClasses];
  if (cls < 0 || cls >= kGroupClasses || kind < 0 || kind >= kGroupKinds) return 0;
  const int i = kind * kGroupClasses + cls;
  std::lock_guard<std::mutex> lk(mu);
  if (!cache[i])
    cache[i] = resident_blocks(kGroupKernels[i].fn, grouped_lds(cls, kind),
                               kind == kGroupShared ? 64 * kDotbWaves : 64 * kGroupWaveWpb);
  return cache[i];
}

hipError_t launch_grouped(int cls, int kind, const GroupProblem *probs,
                          const GroupTile *tiles, int64_t ntiles, int nblocks, hipStream_t st) {
  if (ntiles <= 0) return hipSuccess;
  if (cls < 0 || cls >= kGroupClasses || kind < 0 || kind >= kGroupKinds)
    return hipErrorInvalidValue;
  if (nblocks < 1) nblocks = 1;
  return kGroupKernels[kind * kGroupClasses + cls].launch(dim3((unsigned)nblocks), probs, tiles,
                                                          ntiles, st);
}

}  // namespace hml
