/*
 * hml_ext.h -- extended C-ABI of the MI355X skinny SGEMM/SGEMV library.
 *
 * The reference exposes only the host-pointer entry points of hml_sgemm.h /
 * hml_sgemv.h, plus C++ internals that its apps call directly
 * (include/hml_sgemm_kernel.h:68-101, include/hml_sgemv_kernel.h:72-139).
 * Those internals hand out CUDA __global__ function pointers and cannot be a
 * C ABI; this header replaces them with plain-pointer, status-returning
 * calls.  Every pointer argument named A/B/C/x/y/ws is a DEVICE pointer
 * unless stated otherwise; `stream` is a hipStream_t passed as void* (NULL =
 * the null stream).  Calls are asynchronous with respect to the host unless
 * stated otherwise.
 */
#ifndef HML_EXT_H_INCLUDED_
#define HML_EXT_H_INCLUDED_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference path has none: it exits) ------------- */
typedef enum {
  HML_STATUS_SUCCESS             = 0,
  HML_STATUS_INVALID_VALUE       = 1,
  HML_STATUS_HIP_ERROR           = 2,
  HML_STATUS_NO_KERNEL           = 3,
  HML_STATUS_ALLOC_FAILED        = 4,
  HML_STATUS_IO_ERROR            = 5,
  HML_STATUS_UNSUPPORTED         = 6,
  HML_STATUS_WORKSPACE_TOO_SMALL = 7,
  HML_STATUS_VALIDATION_FAILED   = 8,
  HML_STATUS_PARSE_ERROR         = 9
} hmlStatus;

/* The _ENUM_MIN_ / _ENUM_MAX_ members give the enums the full 32-bit range,
 * so a C caller's out-of-range value reaches the argument checks (and is
 * rejected there) instead of being undefined on the C++ side. */
typedef enum {
  HML_ROW_MAJOR = 0,
  HML_COL_MAJOR = 1,
  HML_ORDER_ENUM_MIN_ = -0x7fffffff - 1,
  HML_ORDER_ENUM_MAX_ = 0x7fffffff
} hmlOrder;
typedef enum {
  HML_OP_N = 0,
  HML_OP_T = 1,
  HML_OP_ENUM_MIN_ = -0x7fffffff - 1,
  HML_OP_ENUM_MAX_ = 0x7fffffff
} hmlOp;

/* flags for the Ex entry points */
#define HML_FLAG_READ_C_ALWAYS 0x1u /* read C (or y) even when beta == 0, as
                                       the reference kernels do; without it
                                       beta == 0 never reads C (BLAS rule). */

/* ---- kernel variants ----------------------------------------------------
 * Replaces HmlSgemmKernelType/Config (include/hml_sgemm_kernel.h:33-43) and
 * HmlSgemvKernelType/Config (include/hml_sgemv_kernel.h:38-49): a variant is
 * a (family, tile parameters, split) tuple instead of a function pointer. */
typedef enum {
  HML_SGEMM_FAMILY_BASIC  = 0, /* fixed default variant                   */
  HML_SGEMM_FAMILY_DOT    = 1, /* MFMA f32, streamed operand contiguous in K */
  HML_SGEMM_FAMILY_AXPY   = 2, /* MFMA f32, streamed operand contiguous in M */
  HML_SGEMM_FAMILY_GEMV   = 3, /* N == 1 routed to the GEMV kernels        */
  HML_SGEMM_FAMILY_BF16X3 = 4, /* split-precision bf16 MFMA, 2-term split  */
  HML_SGEMM_FAMILY_DENSE  = 5, /* plain FP32 MFMA dense tile kernel        */
  HML_SGEMM_FAMILY_BF16X6 = 6, /* split-precision bf16 MFMA, 3-term split  */
  HML_SGEMM_FAMILY_DOTB   = 7, /* MFMA f32, K-contiguous A, B shared by the
                                  workgroup through LDS                   */
  HML_SGEMM_FAMILY_GEMVN  = 8, /* N in {2, 4, 8}, K-contiguous A: the GEMV
                                  DOT kernels with N accumulators per row */
  HML_SGEMM_FAMILY_F16X3  = 9  /* split-precision fp16 MFMA, 2-term split
                                  with power-of-two row / column scales   */
} hmlSgemmFamily;

/* Skinny variants (DOT, AXPY): a wave computes rt row tiles (16 rows DOT,
 * 64 rows AXPY) x ct column tiles of 16 over one of ksplit K chunks.  DOT:
 * the wpb (1, 2, 4 or 8) waves of a workgroup take consecutive chunks of one
 * tile and sum them in LDS; chunks beyond one workgroup are summed by a
 * reduction pass; ct <= 4 and rt * ct <= 8.  DOTB: the 4 waves (wpb = 4)
 * of a workgroup take 4 consecutive row tiles and share each stage of B
 * through LDS; every K chunk is a workgroup and splits are summed by a
 * reduction pass; rt in {1, 2}, ct in {1, 2, 4}.  AXPY: like DOT, the wpb
 * (1, 2, 4, or 8 when rt * ct < 8) waves of a workgroup take consecutive
 * chunks of one tile and sum them in LDS; rt in {1, 2}, ct in {1, 2, 4, 8},
 * not 2 x 8.  GEMVN: rt = lanes per row (16, 32, 64), ct = rows per lane
 * group (1, 2, 4), ksplit = chunks of K, wpb (1, 2, 4, 8) as for the DOT
 * SGEMV variants. */
typedef struct {
  int32_t family; /* hmlSgemmFamily                                          */
  int32_t rt;     /* row tiles per wave                                       */
  int32_t ct;     /* column tiles of 16 per wave                              */
  int32_t ksplit; /* K chunks per tile, >= 1                                  */
  int32_t wpb;    /* waves per workgroup                                      */
} hmlSgemmVariant;

typedef enum {
  HML_SGEMV_FAMILY_BASIC = 0,
  HML_SGEMV_FAMILY_DOT   = 1, /* outputs are dot products of contiguous rows */
  HML_SGEMV_FAMILY_AXPY  = 2  /* outputs are sums of contiguous rows         */
} hmlSgemvFamily;

/* SGEMV variants.  DOT with split > 1 and wpb > 1: the wpb waves of a
 * workgroup take consecutive chunks of the same rows and sum them in LDS
 * (split is rounded up to a multiple of wpb), so only split / wpb > 1
 * leaves partials for a reduction pass.  AXPY: the wpb waves of a
 * workgroup take interleaved rows of one column chunk and add their sums in
 * LDS; split chunks of rows are summed by a reduction pass.  wpb is 1, 2, 4
 * or 8. */
typedef struct {
  int32_t family; /* hmlSgemvFamily                                      */
  int32_t lpr;    /* lanes per row (DOT) / lanes per column chunk (AXPY)  */
  int32_t rb;     /* rows per lane group (DOT) / unroll (AXPY)            */
  int32_t split;  /* chunks of the reduced dimension                      */
  int32_t wpb;    /* waves per workgroup                                  */
} hmlSgemvVariant;

/* ---- GEMM / GEMV on device pointers --------------------------------------
 * C = alpha * op(A) * op(B) + beta * C, C is M x N, op(A) M x K, op(B) K x N,
 * in `order` with leading dimensions lda/ldb/ldc (BLAS convention).
 * workspace may be NULL: the library then uses a buffer it owns per (device,
 * stream), so calls on different streams never share one; it grows on
 * demand with hipMalloc, so such calls must not be captured into a graph
 * the first time a stream needs a larger buffer.  Otherwise ws_bytes must be
 * at least hmlSgemmWorkspaceSize() for the same arguments (for strides past
 * the skinny kernels' 31-bit offsets the call needs the dense kernel's
 * operand copies instead and says how many bytes in its error). */
hmlStatus hmlSgemmEx(hmlOrder order, hmlOp transA, hmlOp transB,
                     int64_t M, int64_t N, int64_t K, float alpha,
                     const float *A, int64_t lda, const float *B, int64_t ldb,
                     float beta, float *C, int64_t ldc, uint32_t flags,
                     void *workspace, size_t ws_bytes, void *stream);

/* hmlSgemmEx for callers that cannot allocate device memory (code under
 * stream capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference, whose every call allocates
 * (src/hml_sgemm.cu:155-210).  Partial sums go to a fixed device array of
 * HML_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of
 * a workspace; the call allocates and frees nothing.  A problem whose
 * variant needs more scratch than that returns
 * HML_STATUS_WORKSPACE_TOO_SMALL.  All calls on one GPU share its array:
 * calls that overlap in time on one GPU race (order them on one stream).
 * Other arguments, selection and results as hmlSgemmEx. */
#define HML_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
hmlStatus hmlSgemmExNoAlloc(hmlOrder order, hmlOp transA, hmlOp transB,
                            int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, float beta, float *C, int64_t ldc,
                            uint32_t flags, void *stream);

/* Same with an explicit variant (autotuner, tests).  variant->family must
 * match the layout (DOT needs op(A) contiguous along K after orientation,
 * AXPY contiguous along M); a mismatched or invalid variant returns
 * HML_STATUS_NO_KERNEL. */
hmlStatus hmlSgemmExVariant(hmlOrder order, hmlOp transA, hmlOp transB,
                            int64_t M, int64_t N, int64_t K, float alpha,
                            const float *A, int64_t lda, const float *B,
                            int64_t ldb, float beta, float *C, int64_t ldc,
                            uint32_t flags, const hmlSgemmVariant *variant,
                            void *workspace, size_t ws_bytes, void *stream);

/* Workspace bytes needed by hmlSgemmEx (variant == NULL: the variant the
 * selector would choose) or by hmlSgemmExVariant. */
size_t hmlSgemmWorkspaceSize(hmlOrder order, hmlOp transA, hmlOp transB,
                             int64_t M, int64_t N, int64_t K,
                             const hmlSgemmVariant *variant);

/* The variant hmlSgemmEx would run for these arguments, and which selector
 * produced it (0 built-in heuristic, 1 config table, 2 learned MLP; the
 * order is given with hmlSelectorInstall). */
hmlStatus hmlSgemmSelect(hmlOrder order, hmlOp transA, hmlOp transB,
                         int64_t M, int64_t N, int64_t K, float beta,
                         hmlSgemmVariant *variant, int32_t *source);

/* Dense SGEMM on the matrix cores, no reference counterpart (the reference
 * has only its skinny kernels).  precision: 0 = plain FP32 MFMA
 * (v_mfma_f32_16x16x4f32), 1 = bf16x3 (FP32 inputs split into hi/lo bf16,
 * products hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16, FP32
 * accumulation), 2 = bf16x6 (hi/mid/lo split, the six products down to
 * 2^-16 relative weight), 3 = f16x3 (rows of A and columns of B scaled by
 * powers of two so their largest magnitude is in [2^14, 2^15), then split
 * into hi/lo fp16, products hi*hi and hi*lo + lo*hi accumulated separately
 * on v_mfma_f32_16x16x32_f16, the scales removed exactly in the epilogue;
 * elements below about 2^-24 of their row's largest lose precision, and a
 * row or column holding Inf or NaN is not scaled).  Other arguments as
 * hmlSgemmEx; the workspace holds the split planes
 * (hmlSgemmDenseWorkspaceSize). */
hmlStatus hmlSgemmDense(int32_t precision, hmlOrder order, hmlOp transA,
                        hmlOp transB, int64_t M, int64_t N, int64_t K,
                        float alpha, const float *A, int64_t lda,
                        const float *B, int64_t ldb, float beta, float *C,
                        int64_t ldc, uint32_t flags, void *workspace,
                        size_t ws_bytes, void *stream);
size_t hmlSgemmDenseWorkspaceSize(int32_t precision, int64_t M, int64_t N,
                                  int64_t K);

/* Grouped skinny SGEMM, no reference counterpart: C_i = alpha * A_i * B_i +
 * beta * C_i for `count` independent ROW-MAJOR problems (M_i >= N_i), given
 * as host arrays of device pointers and extents.  The plan owns the device
 * descriptors and, per tile class (N <= 16, <= 32, <= 48, larger: 1, 2, 3 or
 * 4 column tiles of 16 per pass), a longest-first tile schedule for one
 * persistent kernel.  Run launches one kernel per
 * non-empty class; with use_graph those launches are captured into a
 * hipGraph on the first run and replayed after.  The A/B/C pointers are
 * captured at creation: rerun after changing their contents, create a new
 * plan for new pointers.  HML_STATUS_UNSUPPORTED for N_i > M_i, or strides
 * so large that 31 * lda_i + K_i or (K_i + 16) * ldb_i reaches 2^29
 * elements. */
typedef struct hmlGroupedPlan_t *hmlGroupedPlan;
hmlStatus hmlSgemmGroupedCreate(hmlGroupedPlan *plan, int32_t count,
                                const int64_t *M, const int64_t *N,
                                const int64_t *K, const float *const *A,
                                const int64_t *lda, const float *const *B,
                                const int64_t *ldb, float *const *C,
                                const int64_t *ldc, float alpha, float beta);
hmlStatus hmlSgemmGroupedRun(hmlGroupedPlan plan, int32_t use_graph,
                             void *stream);
hmlStatus hmlSgemmGroupedInfo(hmlGroupedPlan plan, int64_t *ntiles,
                              int32_t *nblocks, size_t *device_bytes);
/* Kernel kind of each of the HML_GROUPED_CLASSES tile classes (kinds[0..3],
 * -1 for a class with no problems): HML_GROUPED_WAVE2 (each wave runs
 * 32-row tiles), HML_GROUPED_WAVE1 (16-row tiles, more resident waves) or
 * HML_GROUPED_SHARED (each workgroup runs a block of 4 row tiles that read
 * every stage of B once, through LDS).  A new plan takes the measured
 * default (shared for N <= 32, wave2 above) unless HML_GROUPED_KERNEL =
 * wave, wave1 or shared names one kind for every class. */
#define HML_GROUPED_CLASSES 4
#define HML_GROUPED_WAVE2 0
#define HML_GROUPED_WAVE1 1
#define HML_GROUPED_SHARED 2
hmlStatus hmlSgemmGroupedKinds(hmlGroupedPlan plan, int32_t *kinds);
/* Sets the kernel kind of one class (waits for the device first). */
hmlStatus hmlSgemmGroupedSetKind(hmlGroupedPlan plan, int32_t cls, int32_t kind);
/* The plan-level autotuner: times every kernel kind of every non-empty
 * class on the plan's own problems (one warm-up and `reps` timed runs on
 * `stream`) and keeps the fastest per class.  ms, if not NULL, receives the
 * mean milliseconds per run at ms[class * 3 + kind] for the 4 classes (-1
 * where not run).  C
 * is saved into a temporary device copy and restored, so it holds what it
 * held before; the plan must not be running meanwhile. */
hmlStatus hmlSgemmGroupedAutotune(hmlGroupedPlan plan, int32_t reps, void *stream, float *ms);
hmlStatus hmlSgemmGroupedDestroy(hmlGroupedPlan plan);

/* y = alpha * op(A) * x + beta * y; A is M x N in `order`; incx/incy must
 * be 1. */
hmlStatus hmlSgemvEx(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                     float alpha, const float *A, int64_t lda, const float *x,
                     int64_t incx, float beta, float *y, int64_t incy,
                     uint32_t flags, void *workspace, size_t ws_bytes,
                     void *stream);

hmlStatus hmlSgemvExVariant(hmlOrder order, hmlOp trans, int64_t M,
                            int64_t N, float alpha, const float *A,
                            int64_t lda, const float *x, int64_t incx,
                            float beta, float *y, int64_t incy,
                            uint32_t flags, const hmlSgemvVariant *variant,
                            void *workspace, size_t ws_bytes, void *stream);

size_t hmlSgemvWorkspaceSize(hmlOrder order, hmlOp trans, int64_t M,
                             int64_t N, const hmlSgemvVariant *variant);

/* The SGEMV variant hmlSgemvEx would run, and its source as for
 * hmlSgemmSelect (the learned net chooses among its GEMV classes). */
hmlStatus hmlSgemvSelect(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                         float beta, hmlSgemvVariant *variant,
                         int32_t *source);

/* ---- config store ----------------------------------------------------------
 * Replaces hmlSgemmKernelConfigReadFile / ConfigPrint
 * (src/hml_sgemm_kernel.cu:557-602) and the SGEMV pair
 * (src/hml_sgemv_kernel.cu:634-679) with status-returning calls.  Loading a
 * v1.0 file keeps the reference's table (K 0..96 x colStops 0..16, N
 * 0..1024); a v2.0 file (MI355X) holds variants keyed by power-of-two shape
 * buckets.  Save writes v2.0 when v2 rules are active, else v1.0. */
hmlStatus hmlSgemmConfigLoad(const char *path);
hmlStatus hmlSgemmConfigSave(const char *path);
hmlStatus hmlSgemvConfigLoad(const char *path);
hmlStatus hmlSgemvConfigSave(const char *path);
void      hmlConfigClear(void);
/* v1 table entry; returns 1 if a v1 table is loaded, else 0. */
int       hmlSgemmConfigV1Get(int K, int colStops, int *type, int *rowStops);
int       hmlSgemvConfigV1Get(int N, int *type, int *blockStops, int *rowStops);
int64_t   hmlSgemmConfigRuleCount(void); /* active v2 rules */
int64_t   hmlSgemvConfigRuleCount(void);

/* ---- exhaustive autotuner -------------------------------------------------
 * Replaces hmlSgemmKernelConfigOnline (src/hml_sgemm_kernel.cu:356-555) and
 * hmlSgemvKernelConfigOnline (src/hml_sgemv_kernel.cu:391-632) for one
 * shape: every candidate variant is validated against a reference output
 * (itself checked by the FP64 kernel), then timed `trials` times with
 * hipEvents (after a cache flush when `flush`), and the median decides.
 * Each (shape, variant, median ms) is appended to the CSV `db_path` (may be
 * NULL).  heuristic_ms receives the built-in heuristic variant's time. */
typedef struct {
  int64_t M, N, K;
  int32_t order, transA, transB, beta0;
} hmlGemmShape;

/* Optional replacement of GPU timing (for host-only tests of the search):
 * returns milliseconds, or < 0 for a failed candidate.  kind 0 = SGEMM
 * (dims = M,N,K; variant = family,rt,ct,ksplit,wpb), 1 = SGEMV (dims =
 * M,N,0; variant = family,lpr,rb,split,wpb).  NULL restores GPU timing. */
typedef double (*hmlAutotuneTimer)(int32_t kind, const int64_t *dims,
                                   const int32_t *variant, void *user);
void    hmlAutotuneSetTimer(hmlAutotuneTimer fn, void *user);

int32_t hmlSgemmCandidates(const hmlGemmShape *shape, hmlSgemmVariant *out,
                           int32_t cap);
int32_t hmlSgemvCandidates(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                           hmlSgemvVariant *out, int32_t cap);
hmlStatus hmlSgemmAutotune(const hmlGemmShape *shape, int32_t trials,
                           int32_t flush, uint64_t seed, hmlSgemmVariant *best,
                           double *best_ms, double *heuristic_ms,
                           const char *db_path);
hmlStatus hmlSgemvAutotune(hmlOrder order, hmlOp trans, int64_t M, int64_t N,
                           int32_t trials, int32_t flush, uint64_t seed,
                           hmlSgemvVariant *best, double *best_ms,
                           double *heuristic_ms, const char *db_path);
/* Summary tables of a timing database, the counterpart of the reference's
 * hmlSgemmKernelConfigPrintStats (src/hml_sgemm_kernel.cu:318-354) and of
 * the tables its SGEMV sweep prints (src/hml_sgemv_kernel.cu:391-632): per
 * layout group and variant family, the best GFLOP/s for every (K, N) (SGEMV:
 * (M, N)) and the max / min spread of the family's valid variants in
 * percent, then the overall winner per cell.  kind 0 = SGEMM rows, 1 =
 * SGEMV rows; out_path NULL writes to stderr.  hmlSgemmAutotune /
 * hmlSgemvAutotune skip candidates that db_path already records for the
 * shape (an interrupted sweep resumes), and hmlSgemmKernelConfig(NULL) /
 * hmlSgemvKernelConfig(NULL) sweep with $HML_AUTOTUNE_DB as their database
 * when it is set (printing these tables at HML_VERBOSE >= 2). */
hmlStatus hmlAutotuneStats(const char *db_path, int32_t kind,
                           const char *out_path);
/* Install (or replace) the v2 rule of the shape's bucket. */
hmlStatus hmlSgemmConfigAddRule(const hmlGemmShape *shape,
                                const hmlSgemmVariant *variant);
hmlStatus hmlSgemvConfigAddRule(hmlOrder order, hmlOp trans, int64_t M,
                                int64_t N, const hmlSgemvVariant *variant);

/* ---- learned kernel selector ---------------------------------------------
 * No reference counterpart (the reference selects from a static table,
 * src/hml_sgemm_kernel.cu:189-237).  A small FP32 MLP (n_in -> n_hidden
 * ReLU layers of `hidden` -> n_out logits) whose training (masked
 * soft-target cross-entropy, Adam) and inference run in HIP kernels.  All
 * array arguments are HOST arrays, row-major [rows x columns]; calls are
 * synchronous.  Weights are laid out per layer as W[in][out] then bias[out]. */
typedef struct hmlMlp_t *hmlMlp;
hmlStatus hmlMlpCreate(hmlMlp *net, int32_t n_in, int32_t hidden,
                       int32_t n_hidden, int32_t n_out, uint64_t seed);
hmlStatus hmlMlpDestroy(hmlMlp net);
int64_t   hmlMlpParamCount(hmlMlp net);
hmlStatus hmlMlpGetParams(hmlMlp net, float *params, int64_t count);
hmlStatus hmlMlpSetParams(hmlMlp net, const float *params, int64_t count);
/* Mean loss and the gradient of the summed-over-batch / batch loss for one
 * batch (X: batch x n_in, T and mask: batch x n_out; mask 1 = class allowed,
 * T sums to 1 over allowed classes). */
hmlStatus hmlMlpLossGrad(hmlMlp net, const float *X, const float *T,
                         const float *mask, int32_t batch, float *loss,
                         float *grads);
/* Shuffled minibatch Adam over `count` rows for `epochs`; epoch_loss (may
 * be NULL) receives the mean loss of each epoch. */
hmlStatus hmlMlpTrain(hmlMlp net, const float *X, const float *T,
                      const float *mask, int64_t count, int32_t epochs,
                      int32_t batch, float lr, uint64_t seed,
                      float *epoch_loss);
/* Masked softmax probabilities (count x n_out). */
hmlStatus hmlMlpPredict(hmlMlp net, const float *X, const float *mask,
                        int64_t count, float *probs);
hmlStatus hmlMlpSave(hmlMlp net, const char *path);
hmlStatus hmlMlpLoad(hmlMlp *net, const char *path);

/* The selector's features (log2 M/16, log2 N/16, log2 K/16, transA,
 * transB, beta==0 of the row-major equivalent problem) and its class space
 * (skinny MFMA variants by row tiles, column tiles, K split and waves per
 * block; dense FP32 tiles; GEMV kernels).  Mask: 1 for classes that can run
 * the shape.  Selection order of hmlSgemm* / hmlSgemv*: a loaded table's
 * rule for the problem's own shape bucket, the installed net, the nearest
 * table rule, the built-in heuristic.  The net shipped beside the library
 * (config/hml_selector_mi355x.bin) is installed on a process's first
 * selection unless HML_SELECTOR is off / 0 (another value names a model
 * file).  Install: the net replaces it (a host copy of its weights picks
 * each first-seen problem on the calling thread, no device work; decisions
 * memoised per problem, at most 4096); NULL uninstalls. */
int32_t   hmlSelectorFeatureCount(void);
int32_t   hmlSelectorClassCount(void);
hmlStatus hmlSelectorFeatures(const hmlGemmShape *shape, float *features);
hmlStatus hmlSelectorMask(const hmlGemmShape *shape, float *mask);
hmlStatus hmlSelectorVariant(const hmlGemmShape *shape, int32_t cls,
                             hmlSgemmVariant *variant);
int32_t   hmlSelectorClassOf(const hmlGemmShape *shape,
                             const hmlSgemmVariant *variant);
hmlStatus hmlSelectorInstall(hmlMlp net);

/* dst[i] = lo + (hi - lo) * u_i with u_i from a counter-based splitmix64
 * stream of `seed` (device-side synthetic inputs). */
hmlStatus hmlFillUniform(float *dst, size_t n, uint64_t seed, float lo,
                         float hi, void *stream);

/* ---- FP64 check kernel ----------------------------------------------------
 * Computes R = alpha * op(A) op(B) + beta * C0 in FP64 on the GPU and
 * compares it with the FP32 result C.  stats (host, 6 doubles):
 *   [0] sum (C - R)^2          [1] sum R^2
 *   [2] sum (|alpha| |op(A)||op(B)| + |beta| |C0|)^2   (for the gamma_K bound)
 *   [3] max |C - R|            [4] max |C - R| / (|alpha||A||B|+|beta||C0|)
 *   [5] count of non-finite C where R is finite
 * C0 may be NULL when beta == 0.  Synchronous. */
hmlStatus hmlSgemmCheckF64(hmlOrder order, hmlOp transA, hmlOp transB,
                           int64_t M, int64_t N, int64_t K, float alpha,
                           const float *A, int64_t lda, const float *B,
                           int64_t ldb, float beta, const float *C0,
                           const float *C, int64_t ldc, double *stats);

/* ---- device / memory helpers (so hosts need no other GPU runtime) -------- */
hmlStatus hmlInit(void);
int       hmlDeviceCount(void);
/* Makes `device` current for the calling thread (the reference apps' -u,
 * apps/gemm/hml_sgemm_main.cu:430). */
hmlStatus hmlSetDevice(int device);
/* name (>= 64 bytes), compute units, L2 bytes, HBM bytes, clock kHz */
hmlStatus hmlDeviceInfo(int device, char *name, int32_t *cus,
                        int64_t *l2_bytes, int64_t *hbm_bytes,
                        int32_t *clock_khz);
hmlStatus hmlDeviceMalloc(void **ptr, size_t bytes);
hmlStatus hmlDeviceFree(void *ptr);
hmlStatus hmlMemcpyHtoD(void *dst, const void *src, size_t bytes);
hmlStatus hmlMemcpyDtoH(void *dst, const void *src, size_t bytes);
hmlStatus hmlMemcpyDtoD(void *dst, const void *src, size_t bytes);
hmlStatus hmlMemsetD8(void *dst, int value, size_t bytes);
hmlStatus hmlDeviceSynchronize(void);
hmlStatus hmlStreamCreate(void **stream);
hmlStatus hmlStreamDestroy(void *stream);
hmlStatus hmlStreamSynchronize(void *stream);
hmlStatus hmlEventCreate(void **event);
hmlStatus hmlEventDestroy(void *event);
hmlStatus hmlEventRecord(void *event, void *stream);
hmlStatus hmlEventSynchronize(void *event);
hmlStatus hmlEventElapsedMs(float *ms, void *start, void *stop);
/* Read-bandwidth probe for benchmarks: streams `bytes` (a multiple of 16)
 * of device memory with 16-byte non-temporal loads and writes nothing (sink,
 * 4 device bytes, is written only in a case that does not occur). */
hmlStatus hmlReadProbe(const void *src, size_t bytes, void *sink, void *stream);
/* Writes `bytes` of junk through the caches so the next timed kernel sees
 * cold L2/MALL (used by the autotuner and benchmarks). */
hmlStatus hmlFlushCaches(void *stream);

const char *hmlStatusString(hmlStatus s);
/* Thread-local text of the last failing call (empty if none). */
const char *hmlLastError(void);

#ifdef __cplusplus
}
#endif

#endif /* HML_EXT_H_INCLUDED_ */
