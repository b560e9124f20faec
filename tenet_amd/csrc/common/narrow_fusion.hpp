// Narrow contractions as one recoupling pass (plan built on the host).
//
// A contraction whose core GEMM jobs all have few contracted (k <= 32) and
// few output (n <= 16) columns is bound by HBM: every element of the large
// lhs and of the output is touched once, with O(k) flops each.  The general
// route (lhs transform -> GEMM -> output transform) reads and writes the
// whole payload in each of its passes.  Fused, each (job, row tree) of the
// core becomes ONE Multi recoupling group
//
//     out_n = sum_k U[n][k] lhs_k,
//     U[n][k] = c_out(n) * c_lhs(k) * c_rhs(k, n) * rhs[k, n]
//
// whose sources are the untransformed lhs blocks of its k columns and whose
// destinations are the final output blocks of its n columns.  The rhs
// entries of U come from the (small) rhs payload through a gather table
// (rhs element, coefficient product) evaluated on the device before the
// pass, so one plan serves every rhs value.  The row trees are the
// row boxes of the Single blocks of the lhs (or output) transform.
//
// Eligible (else the contraction keeps the general route): lhs op identity;
// lhs, rhs and output transforms absent or Single-only, at least one of lhs
// and output present; no inactive lhs layouts; no uncovered core region when
// an output transform exists; every job 1 <= k <= 32, 1 <= n <= 16; every
// transform block a row box x column box of its job's matrix.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/tenet_hip.h"

namespace tnt_fuse {

constexpr int64_t kMaxK = 32;  // sources per Multi group (the recoupling kernel's register bound)
constexpr int64_t kMaxN = 16;  // destinations per Multi group
constexpr int64_t kMaxRhs = int64_t(1) << 24;  // core rhs elements mapped on the host

struct NarrowFusion {
  std::vector<tnt_layout> layouts;
  std::vector<tnt_transform_block> blocks;
  std::vector<double> coeffs;       // placeholder U (zeros); the gather writes it
  std::vector<uint32_t> inactive;   // output layouts written by no group
  std::vector<int64_t> gather_idx;  // per U entry: rhs payload element, -1 = structural zero
  std::vector<double> gather_scale; // per U entry: c_out * c_lhs * c_rhs
  tnt_transform_desc desc{};        // the fused transform (points into the vectors above)
  int64_t rhs_src_len = 0;          // rhs payload elements the gather may read
  std::string reason;               // why the contraction is not eligible
};

// true: `f` holds the fused plan; false: `f.reason` names the failed condition.
bool plan_narrow_fusion(const tnt_contract_desc* d, NarrowFusion& f);

}  // namespace tnt_fuse
