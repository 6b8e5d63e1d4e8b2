// hml_internal.h -- shared declarations between the host library and the
// HIP kernel translation units.  Not installed; not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

namespace hml {

// Streamed-operand forms of the skinny kernels (see hml_k_skinny.hip).
enum Form : int {
  kFormDot = 0,   // A(m,k) at A[m*sAm + k], contiguous along K
  kFormAxpy = 1,  // A(m,k) at A[m + k*sAk], contiguous along M
};

// k-step (K values consumed per loop iteration) of the AXPY form; its
// K-split granularity is a multiple of it.
constexpr int kAxpyKStep = 16;
// Rows covered by one "row tile" of each form.
constexpr int kDotRowsPerTile = 16;
constexpr int kAxpyRowsPerTile = 64;

struct SkinnyParams {
  const float *A;   // streamed operand, A(m,k) = A[m + k*sAk]
  int64_t sAm, sAk; // its strides (sAm == 1)
  const float *B;   // B(k,n) = B[k*sBk + n*sBn]
  int64_t sBk, sBn;
  float *C;         // output
  int64_t sCm, sCn;
  float *P;         // partials [kgroups][M*N] when kgroups > 1
  int64_t M, N, K;
  int64_t kchunk;   // K range per wave, multiple of the k-step
  int32_t kgroups;  // workgroups along K
  int32_t nct_pad;  // column tiles covered by the grid (multiple of the variant's CT)
  float alpha, beta;
  int32_t read_c;   // C participates in the epilogue
};

struct GemvParams {
  const float *A;
  int64_t lda;
  const float *x;
  float *y;
  float *P;         // partials [split][outputs] when split > 1
  int64_t M, N;     // DOT: M outputs, rows of length N; AXPY: M reduced rows, N outputs
  int64_t chunk;    // length of the reduced dimension per split
  int32_t split;
  int32_t wsum;     // DOT: the waves of a workgroup take consecutive chunks of
                    // the same rows and sum them in LDS (split % wpb == 0);
                    // partials [split / wpb][outputs] when split > wpb
  float alpha, beta;
  int32_t read_c;
};
// Most waves per workgroup of the GEMV DOT kernels.
constexpr int kGemvMaxWaves = 8;

// Several right-hand sides on the GEMV DOT structure (k_gemv_multi):
// C(m, n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m, n), n < NV columns.
struct GemvNParams {
  const float *A;   // A(m,k) = A[m*lda + k]
  int64_t lda;
  const float *B;   // B(k,n) = B[k*sBk + n*sBn]
  int64_t sBk, sBn;
  float *C;
  int64_t sCm, sCn;
  float *P;         // partials [slices][M*NV] when the split needs them
  int64_t M, K;
  int64_t chunk;    // K per split
  int32_t split;
  int32_t wsum;     // the waves of a workgroup sum their chunks in LDS
  float alpha, beta;
  int32_t read_c;
};
// Instantiated: lpr in {16, 32, 64}, rb in {1, 2, 4}, nv in {2, 4, 8}, wpb
// in {1, 2, 4, 8}; bvec = B K-contiguous (sBk == 1).
bool gemv_multi_has_variant(int lpr, int rb, int nv);
bool gemv_multi_offsets_fit(int lpr, int rb, int nv, bool bvec, const GemvNParams &p);
hipError_t launch_gemv_multi(int lpr, int rb, int nv, bool bvec, int wpb, const GemvNParams &p,
                             hipStream_t st);

// AXPY-form skinny SGEMM (hml_k_skinny.hip, k_skinny_axpy): B read in
// place, K split over the kw waves of a workgroup and summed there; kgroups
// > 1 splits K further over workgroups (partials in P, summed by
// launch_reduce with R = kgroups).  vec: M % 4 == 0; bvec: sBk == 1.
// Returns hipErrorInvalidValue when (rt, ct, kw) has no instantiation.
hipError_t launch_skinny_axpy(int rt, int ct, bool vec, bool bvec, int kw,
                              const SkinnyParams &p, hipStream_t st);
bool skinny_axpy_runnable(int rt, int ct, int kw);
bool skinny_has_variant(int form, int rt, int ct);

// DOT-form skinny SGEMM (hml_k_skinny.hip, k_skinny_dot): B read in place,
// A staged per wave through LDS, K split over the waves of a workgroup and
// summed there; kgroups > 1 splits K further over workgroups (partials in P,
// summed by launch_reduce with R = kgroups).
constexpr int kDotStageK = 64;  // K per pipeline stage and K-split quantum
struct DotParams {
  const float *A;
  int64_t sAm;                    // A(m,k) = A[m*sAm + k]
  const float *B;
  int64_t sBk, sBn;               // B(k,n) = B[k*sBk + n*sBn]
  float *C;
  int64_t sCm, sCn;
  float *P;                       // [kgroups][M*N] when kgroups > 1
  int64_t M, N, K;
  int64_t kchunk;                 // K per wave, multiple of kDotStageK
  int32_t kgroups;
  int32_t read_c;
  float alpha, beta;
};
// kw = waves per workgroup (each with its own K chunk)
bool skinny_dot_runnable(int rt, int ct, int kw);
// Load kind of the skinny DOT kernels' A (row-major, leading dimension ld
// floats).  Rows that do not start on 128-byte lines share a line between
// the end of one 64-k stage and the start of the next; read
// non-temporally, that line tends to leave L2 between the two stages and
// is fetched twice (the grouped suite's odd-K problems fetched 1.35x their
// bytes that way, 1.07x with cached loads; single problems with odd K ran
// 6-14% faster cached).  Line-aligned rows share nothing and stream faster
// non-temporally (8192 x 16 x 8192: 52 vs 58 us).  With more than one
// column block (ncb), the blocks that share rows run on one XCD at about
// the same time (xcd_block), and cached loads let the later ones read A
// from that XCD's L2.  HML_DOT_A_LOADS = nt / cached forces one kind.
inline bool a_loads_cached(const void *A, int64_t ld, int64_t ncb) {
  static const int force = [] {
    const char *e = getenv("HML_DOT_A_LOADS");
    if (e && !strcmp(e, "nt")) return 0;
    if (e && !strcmp(e, "cached")) return 1;
    return -1;
  }();
  if (force >= 0) return force == 1;
  return ((uintptr_t)A % 128) != 0 || (ld * 4) % 128 != 0 || ncb > 1;
}
// Waves of the (rt, ct) tile kernel with 4-wave workgroups resident on the
// current device at once (0 if unknown).
int skinny_dot_resident_waves(int rt, int ct);
size_t skinny_dot_lds_bytes(int rt, int ct, int kw);
// bvec: B is K-contiguous (sBk == 1), read with 16-byte loads
// k_skinny_dotb: 4 waves on 4 row tiles share each B stage through LDS;
// a K split spans workgroups (kgroups > 1 writes partials).
bool skinny_dotb_runnable(int rt, int ct);
hipError_t launch_skinny_dotb(int rt, int ct, bool bvec, const DotParams &p, hipStream_t st);
hipError_t launch_skinny_dot(int rt, int ct, bool bvec, int kw, const DotParams &p,
                             hipStream_t st);

// out(m,n) = alpha * sum_r P[r*total + m*ninner + n] + beta * out(m,n)
hipError_t launch_reduce(const float *P, int R, int64_t total, int64_t ninner,
                         float *C, int64_t sCm, int64_t sCn, float alpha,
                         float beta, int read_c, hipStream_t st);

// Epilogue only (K == 0, or alpha == 0 under BLAS rules):
// C = alpha*0 + beta*C, or alpha*0 when !read_c.
hipError_t launch_scale(float *C, int64_t M, int64_t N, int64_t sCm,
                        int64_t sCn, float alpha, float beta, int read_c,
                        hipStream_t st);

hipError_t launch_gemv(int family, int lpr, int rb, bool vec, int wpb,
                       const GemvParams &p, hipStream_t st);
bool gemv_has_variant(int family, int lpr, int rb);
// Workgroup sizes (waves) the GEMV kernels take: 1, 2, 4 or 8.
bool gemv_wpb_ok(int family, int wpb);
// Resolves a GEMV variant's split of `reduced` into p->chunk, p->split and
// p->wsum; returns the partial slices the launch writes (1: none, y
// directly; > 1: the split reduction sums that many).
int gemv_resolve_split(int family, int64_t reduced, int lpr, int split, int wpb,
                       GemvParams *p);

struct CheckParams {
  const float *A; int64_t sAm, sAk;
  const float *B; int64_t sBk, sBn;
  const float *C0; const float *C; int64_t sCm, sCn;
  int64_t M, N, K;
  double alpha, beta;
  double *stats;    // [gridDim.x*gridDim.y][6] per-block partial statistics
};
hipError_t launch_check_f64(const CheckParams &p, int64_t *nblocks,
                            hipStream_t st);

hipError_t launch_fill(float *dst, size_t n, float value, hipStream_t st);
// Reads `bytes` (whole 16-byte chunks) and writes nothing (see k_read_probe).
hipError_t launch_read_probe(const void *src, size_t bytes, uint32_t *sink, hipStream_t st);

}  // namespace hml

namespace hml {
// dst[i] = lo + (hi-lo) * u_i, u_i from a counter-based splitmix64 stream.
hipError_t launch_fill_uniform(float *dst, size_t n, uint64_t seed, float lo,
                               float hi, hipStream_t st);
// Per-block statistics of C vs R: [sum (C-R)^2, sum R^2, max|C-R|, max|R|,
// count non-finite C where R is finite], nblocks x 5 doubles.
constexpr int kCompareBlocks = 1024;
hipError_t launch_compare(const float *C, const float *R, int64_t M, int64_t N,
                          int64_t sCm, int64_t sCn, double *stats, int nblocks,
                          hipStream_t st);
}  // namespace hml

namespace hml {
// One problem of a grouped (batched) skinny SGEMM, row-major NN after
// canonicalisation (DOT form): C(m,n) = alpha sum_k A[m*lda+k] B[k*ldb+n] +
// beta C[m*ldc+n].
struct GroupProblem {
  const float *A;
  int64_t lda;
  const float *B;
  int64_t ldb;
  float *C;
  int64_t ldc;
  int64_t M, N, K;
  int32_t read_c;
  float alpha, beta;
};
struct GroupTile {
  int32_t prob;
  int32_t pad;
  int64_t rb;  // row block of grouped_unit_rows() rows
};
// Tile classes of the grouped kernel, by N: column tiles of 16 per pass,
// 0: N <= 16 (one), 1: N <= 32 (two), 2: N <= 48 (three), 3: larger N
// (four, several passes past 64).
constexpr int kGroupClasses = 4;
inline int grouped_class_of(int64_t N) { return N <= 48 ? (int)((N + 15) / 16) - 1 : 3; }
// Kernel kinds of a class: per-wave tiles of 2 or 1 row tiles of 16 rows
// (k_grouped_dot), or workgroup blocks of 4 row tiles that share each stage
// of B through LDS (k_grouped_dotb).
constexpr int kGroupWave2 = 0, kGroupWave1 = 1, kGroupShared = 2;
constexpr int kGroupKinds = 3;
// waves per workgroup of the per-wave kinds (each wave its own entries).
// One-wave workgroups measured within 1% of four-wave ones per class on the
// 512-problem suite (the 4-column class 1.74 against 1.75 ms).
constexpr int kGroupWaveWpb = 4;
// row tiles of 16 rows per wave, and column tiles of 16 per pass
void grouped_kind_shape(int cls, int kind, int *rt, int *ct);
// rows of one work-list entry (GroupTile::rb counts these)
int grouped_unit_rows(int cls, int kind);
// Workgroups of the kernel the device holds at once (0 if unknown).
int grouped_class_blocks(int cls, int kind);
hipError_t launch_grouped(int cls, int kind, const GroupProblem *probs,
                          const GroupTile *tiles, int64_t ntiles, int nblocks, hipStream_t st);
// k-means assignment (hml_k_kmeans.hip).  Screened form: kKmWaves waves
// per workgroup, each on its own 16-row tile; NI = dims / 16 rounded up, NT
// centroid tiles of 16 per screening pass.  Direct form (up to 8 dims, or
// clusters x dims < 1024 up to 32 dims): kKmDirectWaves waves, each on 64
// rows, one per lane; NI = dims.  KC centroids per LDS chunk; tiles = wave tiles of rows.
constexpr int kKmWaves = 8;
constexpr size_t kKmLdsBytes = 163840;      // all of a CU's LDS
constexpr int kKmDirectWaves = 4;
constexpr size_t kKmDirectLdsBytes = 65536; // centroid chunk of the direct form
constexpr int kKmDirectMaxDims = 32;
struct KmeansPlan {
  bool direct = false;
  int ni = 1, nt = 1, KC = 0;
  size_t lds = 0;
  int64_t tiles = 0;
};
// mode: 0 by size, 1 screened, 2 direct
KmeansPlan plan_kmeans(int64_t nrows, int D, int K, int mode);
hipError_t launch_kmeans_assign(const KmeansPlan &pl, uint32_t *assign, const float *rows,
                                int64_t nrows, int D, const float *ctrds, int K, bool fused,
                                hipStream_t st);

}  // namespace hml
