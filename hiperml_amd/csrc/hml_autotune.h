// hml_autotune.h -- exhaustive autotuner (internal entry points).
#pragma once

#include "../../include/hml_ext.h"

namespace hml {

// The reference's hmlSgemmKernelConfigOnline / hmlSgemvKernelConfigOnline
// (src/hml_sgemm_kernel.cu:356-555, src/hml_sgemv_kernel.cu:391-632):
// sweep a fixed shape grid, install the winners as the active table.
hmlStatus autotune_sgemm_online(int verbose);
hmlStatus autotune_sgemv_online(int verbose);

}  // namespace hml
