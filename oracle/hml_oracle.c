/*
 * hml_oracle.c -- CPU oracle for the skinny SGEMM/SGEMV hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  It is used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, always as the checker or as the timed CPU
 * baseline, never as the thing measured on the GPU.
 *
 * Contents
 *   1. Restatement of the reference CPU path (the result the reference's own
 *      apps compare against):
 *        oracle_sgemm_cpu  <- apps/gemm/hml_sgemm_main.cu:260-292 (hmlSgemmCpu)
 *        oracle_sgemv_cpu  <- apps/gemv/hml_sgemv_main.cu:254-283 (hmlSgemvCpu)
 *      Serial loops in the reference order with a float accumulator.  This
 *      file MUST be compiled with -ffp-contract=off (see oracle/Makefile) so
 *      that `sum += a*b` rounds the product and the sum separately, as the
 *      reference's host compiler does on baseline x86-64.
 *   2. Checksums: oracle_csum <- apps/gemm/hml_sgemm_main.cu:294-305
 *                 oracle_ysum <- apps/gemv/hml_sgemv_main.cu:285-292
 *   3. Reference input generator: oracle_hml_random_fill reproduces the
 *      apps' inputs, hmlRandomFloat() = rand()/RAND_MAX (src/hml_utils.cu:60),
 *      filled in the order A, B, C (apps/gemm/hml_sgemm_main.cu:457-470).
 *   4. FP64 reference for any strided layout (the "FP64 oracle"): double
 *      accumulation of the FP32 inputs, plus the |A||B| product used for the
 *      Higham gamma_K error bound.
 *   5. Flop formulas <- include/hml_flops.h:33-45,186,202.
 *   6. k-means nearest-centroid assignment:
 *        oracle_kmeans_assign <- the assign kernels
 *          hmlKmeansAssignCtrdSmem / SmemUnroll / Gmem
 *          (include/hml_kmeans_kernel_template.h:18-88, 90-316, 323-367)
 *          and the CPU path hmlKmeansAssign / hmlKmeansDistance
 *          (src/hml_kmeans_cpu.cu:55-95).
 *      Per row, per centroid k ascending: dist = sum over dims ascending of
 *      delta*delta, delta = x - c; the row takes the first k whose dist is
 *      strictly below every earlier one, starting from FLT_MAX.  `fused`
 *      selects the GPU kernels' arithmetic (nvcc contracts
 *      `dist += delta*delta` into one fused multiply-add by default, and the
 *      reference sets no -fmad flag, CMakeLists.txt) or the CPU path's
 *      (product and sum rounded separately, baseline x86-64).  A row with no
 *      dist below FLT_MAX (NaN or overflowing data) gets numClusts: the CPU
 *      path's value for its first row (it leaves minK over from the previous
 *      row afterwards, src/hml_kmeans_cpu.cu:83-93), the GPU kernels' is
 *      undefined (minK uninitialised).
 *
 * Pinning (what checks each function, besides retyping):
 *   - sgemm/sgemv_cpu: integer-valued inputs give exact results that are
 *     known in closed form (all-ones => K, identity => copy, permutation =>
 *     permuted copy); the float-accumulation order is pinned by a case where
 *     float and double accumulation differ in a known way (1 + 2^-24 * k
 *     sums), and by brute-force agreement with the FP64 oracle within
 *     gamma_K.  See tests/test_oracle.py.
 *   - f64 reference: agrees with numpy float64 matmul (an independent
 *     library routine) on random inputs, and with closed forms.
 *   - csum: exact on integer data; equals math.fsum within 1 ulp on random.
 *   - hml_random_fill: the first values of glibc rand() after srand(1) are a
 *     published constant sequence (1804289383, 846930886, ...).
 *   - kmeans_assign: closed forms (centroids equal to rows: distance 0, the
 *     first equal centroid wins ties), brute-force agreement with a float64
 *     argmin on separated data, and constructed inputs on which the fused
 *     and the separately rounded arithmetic give different assignments whose
 *     expected values follow from exact float64 arithmetic.  See
 *     tests/test_oracle.py.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- 1. reference CPU path restatement ---------------------------------- */

/* C = alpha * A * B + beta * C ; row-major, A M x K, B K x N, C M x N.
 * Follows apps/gemm/hml_sgemm_main.cu:278-287: for each row i, for each
 * column j, sum over e ascending with B accessed down column j. */
void oracle_sgemm_cpu(float *C, const float *A, const float *B,
                      int M, int N, int K, float alpha, float beta) {
  for (int i = 0; i < M; ++i, A += K, C += N) {
    for (int j = 0; j < N; ++j) {
      const float *colB = B + j;
      float sum = 0.0f;
      for (int e = 0; e < K; ++e) {
        float prod = A[e] * colB[(size_t)e * (size_t)N];
        sum = sum + prod;
      }
      float t0 = alpha * sum;
      float t1 = beta * C[j];
      C[j] = t0 + t1;
    }
  }
}

/* y = alpha * A * x + beta * y ; row-major A M x N.
 * Follows apps/gemv/hml_sgemv_main.cu:266-278. */
void oracle_sgemv_cpu(float *y, const float *A, const float *x,
                      int M, int N, float alpha, float beta) {
  for (int i = 0; i < M; ++i, A += N) {
    float ax = 0.0f;
    for (int j = 0; j < N; ++j) {
      float prod = A[j] * x[j];
      ax = ax + prod;
    }
    float t0 = alpha * ax;
    float t1 = beta * y[i];
    y[i] = t0 + t1;
  }
}

/* Row-subsampled variant used only to time the CPU baseline on a bounded
 * sample: computes rows [r0, r1) of the same product. */
void oracle_sgemm_cpu_rows(float *C, const float *A, const float *B,
                           int M, int N, int K, float alpha, float beta,
                           int r0, int r1) {
  (void)M;
  oracle_sgemm_cpu(C + (size_t)r0 * N, A + (size_t)r0 * K, B,
                   r1 - r0, N, K, alpha, beta);
}

void oracle_sgemv_cpu_rows(float *y, const float *A, const float *x,
                           int M, int N, float alpha, float beta,
                           int r0, int r1) {
  (void)M;
  oracle_sgemv_cpu(y + r0, A + (size_t)r0 * N, x, r1 - r0, N, alpha, beta);
}

/* ---- 2. checksums -------------------------------------------------------- */

/* apps/gemm/hml_sgemm_main.cu:294-305: double sum in row-major order. */
double oracle_csum(const float *C, int M, int N) {
  double s = 0.0;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j)
      s += (double)C[(size_t)i * N + j];
  return s;
}

/* apps/gemv/hml_sgemv_main.cu:285-292 */
double oracle_ysum(const float *y, int M) {
  double s = 0.0;
  for (int i = 0; i < M; ++i) s += y[i];
  return s;
}

/* ---- 3. reference input generator --------------------------------------- */

/* hmlRandomFloat() = ((float)rand()) / RAND_MAX (src/hml_utils.cu:60).
 * srand(seed) is applied first when seed_given != 0 (the apps call srand only
 * under -r, apps/gemm/hml_sgemm_main.cu:376-383; glibc's default state equals
 * srand(1)). */
void oracle_hml_random_fill(float *dst, size_t n, unsigned seed, int seed_given) {
  if (seed_given) srand(seed); else srand(1u);
  for (size_t i = 0; i < n; ++i) dst[i] = ((float)rand()) / RAND_MAX;
}

/* Fill A (M*K), B (K*N), C (M*N) in the reference order from one rand()
 * stream, exactly as apps/gemm/hml_sgemm_main.cu:457-470. */
void oracle_hml_gemm_inputs(float *A, float *B, float *C, int M, int N, int K,
                            unsigned seed, int seed_given) {
  if (seed_given) srand(seed); else srand(1u);
  for (size_t i = 0; i < (size_t)M * K; ++i) A[i] = ((float)rand()) / RAND_MAX;
  for (size_t i = 0; i < (size_t)K * N; ++i) B[i] = ((float)rand()) / RAND_MAX;
  for (size_t i = 0; i < (size_t)M * N; ++i) C[i] = ((float)rand()) / RAND_MAX;
}

/* Same for SGEMV: A (M*N), x (N), y (M), apps/gemv/hml_sgemv_main.cu:436-446. */
void oracle_hml_gemv_inputs(float *A, float *x, float *y, int M, int N,
                            unsigned seed, int seed_given) {
  if (seed_given) srand(seed); else srand(1u);
  for (size_t i = 0; i < (size_t)M * N; ++i) A[i] = ((float)rand()) / RAND_MAX;
  for (int j = 0; j < N; ++j) x[j] = ((float)rand()) / RAND_MAX;
  for (int i = 0; i < M; ++i) y[i] = ((float)rand()) / RAND_MAX;
}

/* ---- 4. FP64 oracle ------------------------------------------------------ */

/* C64[m*N+n] = sum_k A(m,k) * B(k,n) in double, with A(m,k) at
 * A[m*sAm + k*sAk] and B(k,n) at B[k*sBk + n*sBn].  Any BLAS layout
 * (row/column-major, transposed) is expressed by the four strides.
 * absC64 (optional) receives sum_k |A(m,k)| |B(k,n)| for error bounds. */
void oracle_gemm_f64(double *C64, double *absC64, const float *A,
                     const float *B, int M, int N, int K,
                     long sAm, long sAk, long sBk, long sBn) {
  for (int m = 0; m < M; ++m) {
    for (int n = 0; n < N; ++n) {
      double s = 0.0, a = 0.0;
      for (int k = 0; k < K; ++k) {
        double x = (double)A[m * sAm + k * sAk];
        double y = (double)B[k * sBk + n * sBn];
        s += x * y;
        a += fabs(x) * fabs(y);
      }
      C64[(size_t)m * N + n] = s;
      if (absC64) absC64[(size_t)m * N + n] = a;
    }
  }
}

/* ---- 5. flop formulas ---------------------------------------------------- */

/* include/hml_flops.h:44-45,202 : FMULS+FADDS = 2MNK */
double oracle_flops_sgemm(double M, double N, double K) { return 2.0 * M * N * K; }
/* include/hml_flops.h:33-34,186 : (MN + 2M) + MN */
double oracle_flops_sgemv(double M, double N) { return 2.0 * M * N + 2.0 * M; }

/* ---- 6. k-means assignment ---------------------------------------------- */

/* rows: numRows x numDims, ctrds: numClusts x numDims, both row-major. */
void oracle_kmeans_assign(uint32_t *out, const float *rows, long numRows,
                          int numDims, const float *ctrds, int numClusts,
                          int fused) {
  for (long r = 0; r < numRows; ++r) {
    const float *x = rows + r * (long)numDims;
    float minDist = FLT_MAX;
    uint32_t minK = (uint32_t)numClusts;
    for (int k = 0; k < numClusts; ++k) {
      const float *c = ctrds + (long)k * numDims;
      float dist = 0.0f;
      for (int d = 0; d < numDims; ++d) {
        const float delta = x[d] - c[d];
        if (fused) {
          dist = fmaf(delta, delta, dist);
        } else {
          const float sq = delta * delta;
          dist = dist + sq;
        }
      }
      if (dist < minDist) {
        minDist = dist;
        minK = (uint32_t)k;
      }
    }
    out[r] = minK;
  }
}
