/*
 * hml_sgemv.h -- single-precision GEMV, reference-compatible entry points.
 * Drop-in for the reference's include/hml_sgemv.h:12-38.
 *
 * y = alpha * A * x + beta * y, A is M x N ROW-MAJOR without padding
 * (src/hml_sgemv.cu:6-11, apps/gemv/hml_sgemv_main.cu:440-446), x has N
 * entries, y has M.  Host pointers; synchronous; y is read even when
 * beta == 0 (include/hml_sgemv_kernel_template.h:103); errors exit(1)
 * (src/hml_sgemv.cu:189-199).
 */
#ifndef HML_SGEMV_H_INCLUDED_
#define HML_SGEMV_H_INCLUDED_

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces hmlSgemvKernelConfig (include/hml_sgemv.h:13,
 * src/hml_sgemv.cu:215-228): loads a .cfg file ("N : kernelType blockStops
 * rowStops", v1.0, or the MI355X v2.0 extension); NULL runs the online
 * exhaustive autotuner. */
void hmlSgemvKernelConfig(char *sgemvKernelConfigFileName);

/* Replaces hmlSgemvBasic (include/hml_sgemv.h:19-25,
 * src/hml_sgemv.cu:38-118). */
void hmlSgemvBasic(float       *y,
                   const float *A,
                   const float *x,
                   const int    M,
                   const int    N,
                   const float  alpha,
                   const float  beta);

/* Replaces hmlSgemv (include/hml_sgemv.h:31-37, src/hml_sgemv.cu:125-213). */
void hmlSgemv(float       *y,
              const float *A,
              const float *x,
              const int    M,
              const int    N,
              const float  alpha,
              const float  beta);

#ifdef __cplusplus
}
#endif

#endif /* HML_SGEMV_H_INCLUDED_ */
