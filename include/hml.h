/*
 * hml.h -- umbrella header of the MI355X skinny SGEMM/SGEMV library.
 * Replaces include/hml.h:9-14 of the reference, which also pulled in the
 * graph-analytics headers; here: the GEMM/GEMV path and the k-means
 * assignment that consumes it.
 */
#ifndef HML_H_INCLUDED_
#define HML_H_INCLUDED_

#include "hml_sgemm.h"
#include "hml_sgemv.h"
#include "hml_ext.h"
#include "hml_kmeans.h"

#endif /* HML_H_INCLUDED_ */
