/*
 * hml_sgemm.h -- single-precision skinny GEMM, reference-compatible entry
 * points.  Drop-in for the reference's include/hml_sgemm.h:14-42.
 *
 * Semantics kept from the reference:
 *   C = alpha * A * B + beta * C, ROW-MAJOR, no padding, no transpose
 *   (src/hml_sgemm.cu:10-11): A is M x K with A[i*K+k], B is K x N with
 *   B[k*N+j], C is M x N with C[i*N+j].  Pointers are HOST pointers; the call
 *   copies to the device, runs, and copies C back before returning
 *   (src/hml_sgemm.cu:155-210).  C is read even when beta == 0, exactly as the
 *   reference kernels do (include/hml_sgemm_kernel_template.h:187).  Errors
 *   print to stderr and exit(1), as the reference does
 *   (src/hml_sgemm.cu:189-199).
 * Differences: declared extern "C" (the reference is C++-mangled), 64-bit
 * offsets (no 2^27-element texture limit, src/hml_sgemm.cu:146-149).
 */
#ifndef HML_SGEMM_H_INCLUDED_
#define HML_SGEMM_H_INCLUDED_

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces hmlSgemmKernelConfig (include/hml_sgemm.h:15,
 * src/hml_sgemm.cu:213-226): loads a kernel-config file (.cfg, v1.0 grammar
 * "K colStops : kernelType rowStops", or the MI355X v2.0 extension of it);
 * NULL runs the online exhaustive autotuner instead. */
void hmlSgemmKernelConfig(char *sgemmKernelConfigFileName);

/* Replaces hmlSgemmBasic (include/hml_sgemm.h:21-28,
 * src/hml_sgemm.cu:39-113): fixed basic kernel, no selection. */
void hmlSgemmBasic(float       *C,
                   const float *A,
                   const float *B,
                   const int    M,
                   const int    N,
                   const int    K,
                   const float  alpha,
                   const float  beta);

/* Replaces hmlSgemm (include/hml_sgemm.h:34-41, src/hml_sgemm.cu:120-211):
 * the "smart" entry, kernel chosen by the config store / learned selector. */
void hmlSgemm(float       *C,
              const float *A,
              const float *B,
              const int    M,
              const int    N,
              const int    K,
              const float  alpha,
              const float  beta);

#ifdef __cplusplus
}
#endif

#endif /* HML_SGEMM_H_INCLUDED_ */
