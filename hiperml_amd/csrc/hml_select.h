// hml_select.h -- canonical problem descriptions and the variant selectors
// (built-in heuristic, config store, learned MLP) shared by the GEMM/GEMV
// dispatch, the autotuner and the config store.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/hml_ext.h"

namespace hml {

// A GEMM after layout canonicalisation and orientation:
// C(m,n) = sum_k A(m,k) B(k,n); A is the streamed (larger) operand.
struct GemmCanon {
  const float *A; int64_t sAm, sAk;
  const float *B; int64_t sBk, sBn;
  float *C; int64_t sCm, sCn;
  int64_t M, N, K;
  int form;      // kFormDot / kFormAxpy
  bool swapped;  // m/n roles exchanged relative to the caller's C
};

// Shape key used by every selector.
struct GemmKey {
  int form;      // 0 dot, 1 axpy (after orientation)
  int64_t M, N, K;
  int beta0;     // beta == 0 and C is not read
  int transA, transB, order;  // caller's layout (features of the learned selector)
  bool gemv_ok;  // N == 1 and x / y are unit-stride, GEMV kernels apply
  int64_t cM, cN;  // caller's M and N (before orientation)
};

struct GemvKey {
  int family;    // 1 DOT, 2 AXPY
  int64_t outputs, reduced;
  int beta0;
};

// Layout -> strides, orientation (larger operand streamed), form.
GemmCanon canonicalize(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M,
                       int64_t N, int64_t K, const float *A, int64_t lda,
                       const float *B, int64_t ldb, float *C, int64_t ldc);
GemmKey make_key(const GemmCanon &g, float beta, uint32_t flags,
                 hmlOrder order, hmlOp ta, hmlOp tb);
GemmCanon dense_canon(hmlOrder order, hmlOp ta, hmlOp tb, int64_t M,
                      int64_t N, int64_t K);
// SGEMV layout -> family and extents.
GemvKey gemv_key(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                 int beta0);
int ilog2_floor(int64_t x);

void heuristic_sgemm(const GemmKey &k, hmlSgemmVariant *v);

// Dense families (plain FP32 / bf16x3 / bf16x6 / f16x3 MFMA tiles).
inline bool is_dense_family(int f) {
  return f == HML_SGEMM_FAMILY_DENSE || f == HML_SGEMM_FAMILY_BF16X3 ||
         f == HML_SGEMM_FAMILY_BF16X6 || f == HML_SGEMM_FAMILY_F16X3;
}
inline int dense_precision(int f) {
  return f == HML_SGEMM_FAMILY_BF16X3 ? 1 : f == HML_SGEMM_FAMILY_BF16X6 ? 2
         : f == HML_SGEMM_FAMILY_F16X3 ? 3 : 0;
}
size_t dense_workspace(int prec, int64_t M, int64_t N, int64_t K);
hmlStatus run_dense(int prec, const float *A, int64_t sAm, int64_t sAk,
                    const float *B, int64_t sBk, int64_t sBn, float *C,
                    int64_t sCm, int64_t sCn, int64_t M, int64_t N, int64_t K,
                    float alpha, float beta, int read_c, void *ws,
                    size_t ws_bytes, hipStream_t st);
void heuristic_sgemv(const GemvKey &k, hmlSgemvVariant *v);
// Row split of a GEMV geometry (lanes per row / column chunk, rows per
// lane group) for this key: the heuristic's rule for any geometry.
int gemv_split_rule(const GemvKey &k, int lpr, int rb);
void basic_sgemm(const GemmKey &k, hmlSgemmVariant *v);
void basic_sgemv(const GemvKey &k, hmlSgemvVariant *v);

// Normalises a variant so it can run (ct divides the padded tile count,
// ksplit at least 1, ...).  Returns false when the variant cannot run this
// key at all (wrong family for the form, no instantiation).
bool sgemm_variant_runnable(const GemmKey &k, const hmlSgemmVariant &v);
bool sgemv_variant_runnable(const GemvKey &k, const hmlSgemvVariant &v);

// Config-store / learned selection; return false when they have no answer.
// exact_only: a v2 table answers only for a rule of the key's own shape
// bucket (a loaded v1 table always answers, as in the reference).
bool config_select_sgemm(const GemmKey &k, hmlSgemmVariant *v, int *source,
                         bool exact_only = false);
bool config_select_sgemv(const GemvKey &k, hmlSgemvVariant *v, int *source,
                         bool exact_only = false);

// Learned MLP selection (when a trained net is installed).  gemv_only:
// choose among the GEMV kernel classes only (SGEMV entry points).
bool learned_select_sgemm(const GemmKey &k, hmlSgemmVariant *v, bool gemv_only = false);
bool learned_select_sgemv(const GemvKey &k, hmlSgemvVariant *v);
// Selector class -> variant for a key (false: the class cannot run it).
bool selector_variant(const GemmKey &k, int cls, hmlSgemmVariant *v);

// Selection entry used by the dispatch: a table rule of the problem's own
// shape bucket (the exhaustive autotuner's answer), else the learned
// selector, else the nearest table rule, else the heuristic.
void select_sgemm(const GemmKey &k, hmlSgemmVariant *v, int *source);
void select_sgemv(const GemvKey &k, hmlSgemvVariant *v, int *source);

}  // namespace hml
