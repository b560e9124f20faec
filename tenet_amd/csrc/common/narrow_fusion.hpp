// Narrow contractions as one recoupling pass (plan built on the host).
//
// A contraction whose core GEMM jobs all have few contracted (k <= 32) and
// few output (n <= 16) columns is bound by HBM: every element of the large
// lhs and of the output is touched once, with O(k) flops each.  The general
// route (lhs transform -> GEMM -> output transform) reads and writes the
// whole payload in each of its passes.  Fused, the contraction is ONE
// recoupling pass from the lhs payload to the output payload made of Multi
// groups
//
//     out_l = sum_s U[l][s] src_s,
//
// one group per set of core row trees that the transforms' Multi blocks mix
// (a row tree of a job alone when both transforms are Single-only): its
// sources are the lhs payload pieces of those rows, its destinations the
// output pieces written from them, and every U entry is a sum of terms
// c_out * c_lhs * c_rhs * rhs[e] over the core columns and the rhs
// transform's terms.  The terms are gathered from the (small) rhs payload on
// the device before the pass (U[t] = sum over gather_ptr[t] <= p <
// gather_ptr[t+1] of scale[p] * rhs[idx[p]]), so one plan serves every rhs
// value.  A Single-only contraction gets one group per (job, row tree) with
// its k sources in column order and one term per U entry,
// U[n][k] = c_out(n) c_lhs(k) c_rhs(k, n) rhs[k, n].
//
// Eligible (else the contraction keeps the general route): lhs op identity;
// at least one of the lhs and output transforms present; no inactive lhs
// layouts; no uncovered core region when an output transform exists; every
// job 1 <= k <= 32, 1 <= n <= 16; every core-side block layout a row box x
// column box of its job's matrix; every group at most 32 sources of one
// row-box shape.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/tenet_hip.h"

namespace tnt_fuse {

constexpr int64_t kMaxK = 32;  // contracted columns per job, and sources per group (the recoupling kernel's bound)
constexpr int64_t kMaxN = 16;  // output columns per job
constexpr int64_t kMaxRhs = int64_t(1) << 24;  // core rhs elements mapped on the host

struct NarrowFusion {
  std::vector<tnt_layout> layouts;
  std::vector<tnt_transform_block> blocks;
  std::vector<double> coeffs;       // placeholder U (zeros); the gather writes it
  std::vector<uint32_t> inactive;   // output layouts written by no group
  std::vector<int64_t> gather_ptr;   // U entry t: terms [gather_ptr[t], gather_ptr[t + 1])
  std::vector<int64_t> gather_idx;   // per term: rhs payload element
  std::vector<double> gather_scale;  // per term: c_out * c_lhs * c_rhs
  tnt_transform_desc desc{};        // the fused transform (points into the vectors above)
  int64_t rhs_src_len = 0;          // rhs payload elements the gather may read
  std::string reason;               // why the contraction is not eligible
};

// true: `f` holds the fused plan; false: `f.reason` names the failed condition.
bool plan_narrow_fusion(const tnt_contract_desc* d, NarrowFusion& f);

}  // namespace tnt_fuse
