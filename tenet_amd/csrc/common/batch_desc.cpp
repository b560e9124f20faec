// Stacked (batched) contraction descriptors; see batch_desc.hpp.
#include "batch_desc.hpp"

namespace tnt_batch {

void OwnedTransform::bind() {
  desc.n_layouts = layouts.size();
  desc.layouts = layouts.data();
  desc.n_blocks = blocks.size();
  desc.blocks = blocks.data();
  desc.n_coeffs = coeffs.size();
  desc.coeffs = coeffs.data();
  desc.n_inactive = inactive.size();
  desc.inactive = inactive.data();
}

namespace {

// One copy of a transform's block table per instance: every layout a block
// (or the inactive list) references is duplicated with its offset moved by
// c * dst_len (destination layouts) or c * src_len (source layouts), so the
// roles of the layouts need not be known up front.  Coefficients are shared.
bool replicate(const tnt_transform_desc& t, size_t copies, OwnedTransform& o, std::string& err) {
  o.coeffs.assign(t.coeffs, t.coeffs + t.n_coeffs);
  auto copy_layout = [&](uint32_t idx, int64_t shift) -> uint32_t {
    tnt_layout l = t.layouts[idx];
    l.offset += shift;
    o.layouts.push_back(l);
    return (uint32_t)(o.layouts.size() - 1);
  };
  for (size_t b = 0; b < t.n_blocks; ++b) {
    const tnt_transform_block& x = t.blocks[b];
    if (x.dst_layout + (uint64_t)x.dst_count > t.n_layouts || x.src_layout + (uint64_t)x.src_count > t.n_layouts) {
      err = "transform block references a layout outside the table";
      return false;
    }
  }
  for (size_t i = 0; i < t.n_inactive; ++i)
    if (t.inactive[i] >= t.n_layouts) {
      err = "inactive layout outside the table";
      return false;
    }
  for (size_t c = 0; c < copies; ++c) {
    const int64_t ds = (int64_t)c * t.dst_len, ss = (int64_t)c * t.src_len;
    for (size_t b = 0; b < t.n_blocks; ++b) {
      tnt_transform_block x = t.blocks[b];
      const uint32_t d0 = (uint32_t)o.layouts.size();
      for (uint32_t i = 0; i < x.dst_count; ++i) copy_layout(x.dst_layout + i, ds);
      const uint32_t s0 = (uint32_t)o.layouts.size();
      for (uint32_t i = 0; i < x.src_count; ++i) copy_layout(x.src_layout + i, ss);
      x.dst_layout = d0;
      x.src_layout = s0;
      o.blocks.push_back(x);
    }
    for (size_t i = 0; i < t.n_inactive; ++i) o.inactive.push_back(copy_layout(t.inactive[i], ds));
  }
  o.bind();
  o.desc.dst_len = t.dst_len * (int64_t)copies;
  o.desc.src_len = t.src_len * (int64_t)copies;
  return true;
}

void keep(const tnt_transform_desc& t, OwnedTransform& o) {
  o.layouts.assign(t.layouts, t.layouts + t.n_layouts);
  o.blocks.assign(t.blocks, t.blocks + t.n_blocks);
  o.coeffs.assign(t.coeffs, t.coeffs + t.n_coeffs);
  o.inactive.assign(t.inactive, t.inactive + t.n_inactive);
  o.bind();
  o.desc.dst_len = t.dst_len;
  o.desc.src_len = t.src_len;
}

}  // namespace

bool batch_contract(const tnt_contract_desc& d, size_t copies, int shared, BatchedContract& out, std::string& err) {
  if (copies < 1) {
    err = "copies must be >= 1";
    return false;
  }
  if (shared != SHARE_LHS && shared != SHARE_RHS && shared != SHARE_NONE) {
    err = "unknown shared operand";
    return false;
  }
  const bool lhs_stacked = shared != SHARE_LHS, rhs_stacked = shared != SHARE_RHS;
  out = BatchedContract{};
  tnt_contract_desc& o = out.desc;
  o = d;
  if (d.lhs) {
    if (lhs_stacked ? !replicate(*d.lhs, copies, out.lhs, err) : (keep(*d.lhs, out.lhs), false)) return false;
    o.lhs = &out.lhs.desc;
  }
  if (d.rhs) {
    if (rhs_stacked ? !replicate(*d.rhs, copies, out.rhs, err) : (keep(*d.rhs, out.rhs), false)) return false;
    o.rhs = &out.rhs.desc;
  }
  if (d.out) {
    if (!replicate(*d.out, copies, out.out, err)) return false;
    o.out = &out.out.desc;
  }
  const int64_t n = (int64_t)copies;
  o.lhs_len = lhs_stacked ? d.lhs_len * n : d.lhs_len;
  o.rhs_len = rhs_stacked ? d.rhs_len * n : d.rhs_len;
  o.core_len = d.core_len * n;
  for (size_t c = 0; c < copies; ++c) {
    for (size_t i = 0; i < d.n_jobs; ++i) {
      tnt_gemm_job j = d.jobs[i];
      j.dst_offset += c * (size_t)d.core_len;
      if (lhs_stacked) j.lhs_offset += c * (size_t)d.lhs_len;
      if (rhs_stacked) j.rhs_offset += c * (size_t)d.rhs_len;
      out.jobs.push_back(j);
    }
    for (size_t i = 0; i < d.n_uncovered; ++i) {
      tnt_region r = d.uncovered[i];
      r.offset += (int64_t)c * d.core_len;
      out.uncovered.push_back(r);
    }
  }
  o.n_jobs = out.jobs.size();
  o.jobs = out.jobs.data();
  o.n_uncovered = out.uncovered.size();
  o.uncovered = out.uncovered.data();
  return true;
}

}  // namespace tnt_batch
