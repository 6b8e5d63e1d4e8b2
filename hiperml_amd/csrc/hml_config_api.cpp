// hml_config_api.cpp -- C ABI of the config store, including the
// reference-compatible hmlSgemmKernelConfig / hmlSgemvKernelConfig.
#include <stdio.h>
#include <stdlib.h>

#include "hml_autotune.h"
#include "hml_config.h"
#include "hml_runtime.h"

using namespace hml;

extern "C" {

hmlStatus hmlSgemmConfigLoad(const char *path) {
  clear_error();
  if (!path) return fail(HML_STATUS_INVALID_VALUE, "path is NULL");
  return config_load_sgemm(path);
}

hmlStatus hmlSgemvConfigLoad(const char *path) {
  clear_error();
  if (!path) return fail(HML_STATUS_INVALID_VALUE, "path is NULL");
  return config_load_sgemv(path);
}

hmlStatus hmlSgemmConfigSave(const char *path) {
  clear_error();
  if (!path) return fail(HML_STATUS_INVALID_VALUE, "path is NULL");
  return config_save_sgemm(path);
}

hmlStatus hmlSgemvConfigSave(const char *path) {
  clear_error();
  if (!path) return fail(HML_STATUS_INVALID_VALUE, "path is NULL");
  return config_save_sgemv(path);
}

void hmlConfigClear(void) { config_clear(); }

int hmlSgemmConfigV1Get(int K, int colStops, int *type, int *rowStops) {
  int t = 0, r = 0;
  int loaded = config_sgemm_v1_get(K, colStops, &t, &r);
  if (type) *type = t;
  if (rowStops) *rowStops = r;
  return loaded;
}

int hmlSgemvConfigV1Get(int N, int *type, int *blockStops, int *rowStops) {
  int t = 0, b = 0, r = 0;
  int loaded = config_sgemv_v1_get(N, &t, &b, &r);
  if (type) *type = t;
  if (blockStops) *blockStops = b;
  if (rowStops) *rowStops = r;
  return loaded;
}

int64_t hmlSgemmConfigRuleCount(void) { return (int64_t)config_sgemm_v2_count(); }
int64_t hmlSgemvConfigRuleCount(void) { return (int64_t)config_sgemv_v2_count(); }

// src/hml_sgemm.cu:213-226: a file name loads the table (openFile exits on
// a missing file, src/hml_utils.cu:65-72); NULL runs the online autotuner.
void hmlSgemmKernelConfig(char *sgemmKernelConfigFileName) {
  hmlStatus s;
  if (sgemmKernelConfigFileName) {
    s = config_load_sgemm(sgemmKernelConfigFileName);
    if (s == HML_STATUS_IO_ERROR) {
      fprintf(stderr, "; Error: Cannot open file '%s'\n", sgemmKernelConfigFileName);
      exit(EXIT_FAILURE);
    }
  } else {
    s = autotune_sgemm_online(verbosity() >= 2 ? 1 : 0);
  }
  if (s != HML_STATUS_SUCCESS) compat_die("hmlSgemmKernelConfig", s);
}

// src/hml_sgemv.cu:215-228
void hmlSgemvKernelConfig(char *sgemvKernelConfigFileName) {
  hmlStatus s;
  if (sgemvKernelConfigFileName) {
    s = config_load_sgemv(sgemvKernelConfigFileName);
    if (s == HML_STATUS_IO_ERROR) {
      fprintf(stderr, "; Error: Cannot open file '%s'\n", sgemvKernelConfigFileName);
      exit(EXIT_FAILURE);
    }
  } else {
    s = autotune_sgemv_online(verbosity() >= 2 ? 1 : 0);
  }
  if (s != HML_STATUS_SUCCESS) compat_die("hmlSgemvKernelConfig", s);
}

}  // extern "C"
