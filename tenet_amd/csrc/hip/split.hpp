// Split (Ozaki-style) GEMM paths, BF16 and INT8 digits: declarations shared with backend.hip.
#pragma once

#include <string>
#include <vector>

#include "kernels.hpp"

namespace tnt_dev {

struct SplitWorkspace;

enum SplitScheme { SPLIT_SCHEME_BF16 = 0, SPLIT_SCHEME_I8 = 1 };

// Runs dst = alpha * A B + beta * dst for every job of the table with the
// split scheme given (BF16 or INT8 digit planes, `slices` planes per operand,
// 2..6).  While `graphs_alive`, memory the workspace outgrows or drops is
// retired instead of freed (a graph may have recorded it).  Returns a TNT
// status; `err` receives the message.
// With `side` (INT8 scheme): the pre-passes of operand side_op (0 = a, 1 = b)
// run on `side` (after the work already queued there, e.g. the transform that
// writes that operand), then `side_done` is recorded there and `s` waits for
// it before the GEMM; the other operand's pre-passes run on `s`.  The caller
// still joins `side_done` into `s` after the call (the call may return before
// it records it).
int split_gemm_run(SplitWorkspace* ws, int scheme, int slices, const std::vector<GJob>& host_jobs,
                   const GJob* dev_jobs, uintptr_t d, uintptr_t a, uintptr_t b, double alpha, double beta,
                   hipStream_t s, bool capturing, bool graphs_alive, std::string& err, hipStream_t side = nullptr,
                   hipEvent_t side_done = nullptr, int side_op = 1);
SplitWorkspace* split_workspace_create();
void split_workspace_destroy(SplitWorkspace* ws);
// Frees the retired memory (call once no graph of the context is alive).
void split_workspace_release(SplitWorkspace* ws);

struct SplitWorkspaceDeleter {
  void operator()(SplitWorkspace* w) const { split_workspace_destroy(w); }
};

}  // namespace tnt_dev
