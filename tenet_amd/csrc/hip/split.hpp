// Split (Ozaki-style) GEMM paths, BF16 and INT8 digits: declarations shared with backend.hip.
#pragma once

#include <string>
#include <vector>

#include "kernels.hpp"

namespace tnt_dev {

struct SplitWorkspace;

enum SplitScheme { SPLIT_SCHEME_BF16 = 0, SPLIT_SCHEME_I8 = 1 };

// Runs dst = alpha * A B + beta * dst for every job of the table with the
// split scheme given (BF16 or INT8 digit planes).  Returns a TNT status; `err`
// receives the message.
int split_gemm_run(SplitWorkspace* ws, int scheme, const std::vector<GJob>& host_jobs, const GJob* dev_jobs,
                   const std::vector<GTile>& host_tiles, const GTile* dev_tiles, uintptr_t d,
                   uintptr_t a, uintptr_t b, double alpha, double beta, hipStream_t s,
                   bool capturing, std::string& err);
int split_set_slices(int slices);
SplitWorkspace* split_workspace_create();
void split_workspace_destroy(SplitWorkspace* ws);

struct SplitWorkspaceDeleter {
  void operator()(SplitWorkspace* w) const { split_workspace_destroy(w); }
};

}  // namespace tnt_dev
