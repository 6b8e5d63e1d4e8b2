// hml_config.h -- config-store types and accessors (internal).
#pragma once

#include <stddef.h>

#include <vector>

#include "../../include/hml_ext.h"

namespace hml {

// Table extents of the v1.0 format (include/hml_sgemm_kernel.h:28-30,
// include/hml_sgemv_kernel.h:28-30).
constexpr int kV1MaxSkinnyK = 96;
constexpr int kV1MaxSkinnyN_Gemm = 96;
constexpr int kV1MaxStops = 16;
constexpr int kV1MaxSkinnyN = 1024;

struct SgemmRuleV1 { int type, rowStops; };
struct SgemvRuleV1 { int type, blockStops, rowStops; };

struct SgemmRuleV2 {
  int form, mlog, nlog, klog;
  hmlSgemmVariant v;
};
struct SgemvRuleV2 {
  int family, olog, rlog;
  hmlSgemvVariant v;
};

hmlStatus config_load_sgemm(const char *path);
hmlStatus config_load_sgemv(const char *path);
hmlStatus config_save_sgemm(const char *path);
hmlStatus config_save_sgemv(const char *path);
void config_clear();
void config_set_sgemm_v2(const std::vector<SgemmRuleV2> &rules);
void config_set_sgemv_v2(const std::vector<SgemvRuleV2> &rules);
// Adds a rule, replacing one with the same bucket key.
void config_add_sgemm_v2(const SgemmRuleV2 &r);
void config_add_sgemv_v2(const SgemvRuleV2 &r);
bool sgemm_config_done();
bool sgemv_config_done();
int config_sgemm_v1_get(int K, int colStops, int *type, int *rowStops);
int config_sgemv_v1_get(int N, int *type, int *blockStops, int *rowStops);
size_t config_sgemm_v2_count();
size_t config_sgemv_v2_count();

}  // namespace hml
