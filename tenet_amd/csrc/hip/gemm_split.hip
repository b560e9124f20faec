// Split-BF16 GEMM (placeholder until the Ozaki kernels land).
#include "../../../include/tenet_hip.h"
#include "split.hpp"

namespace tnt_dev {

struct SplitWorkspace {
  int unused = 0;
};

SplitWorkspace* split_workspace_create() { return new SplitWorkspace(); }
void split_workspace_destroy(SplitWorkspace* ws) { delete ws; }

int split_gemm_run(SplitWorkspace*, const std::vector<GJob>&, const GJob*, const std::vector<GTile>&,
                   const GTile*, uintptr_t, uintptr_t, uintptr_t, double, double, hipStream_t, bool,
                   std::string& err) {
  err = "split-BF16 precision is not available in this build";
  return TNT_E_UNSUPPORTED;
}

}  // namespace tnt_dev
