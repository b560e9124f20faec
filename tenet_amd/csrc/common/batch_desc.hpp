// Batched contraction descriptors: `copies` independent instances of one
// contraction stacked in one artifact, so that one tnt_contract call runs
// every instance with the launches of one (one grouped GEMM table, one pass
// per recoupling kernel class).  Pure descriptor arithmetic, shared by the
// host library and its tests.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../../include/tenet_hip.h"

namespace tnt_batch {

// Which operand all copies share (its payload is given once and its transform
// runs once); the other operand and the destination are stacked: copy c of a
// stacked payload of length L occupies [c L, (c + 1) L).
enum Shared { SHARE_LHS = 0, SHARE_RHS = 1, SHARE_NONE = 2 };

struct OwnedTransform {
  std::vector<tnt_layout> layouts;
  std::vector<tnt_transform_block> blocks;
  std::vector<double> coeffs;
  std::vector<uint32_t> inactive;
  tnt_transform_desc desc{};
  void bind();  // points desc at the vectors
};

struct BatchedContract {
  OwnedTransform lhs, rhs, out;
  std::vector<tnt_gemm_job> jobs;
  std::vector<tnt_region> uncovered;
  tnt_contract_desc desc{};
};

// Builds the stacked descriptor of `copies` (>= 1) instances of d; returns
// false with `err` set when the request is malformed.
bool batch_contract(const tnt_contract_desc& d, size_t copies, int shared, BatchedContract& out, std::string& err);

}  // namespace tnt_batch
