// Host planner of the fused narrow-contraction route (see narrow_fusion.hpp).
#include "narrow_fusion.hpp"

#include <algorithm>
#include <map>
#include <utility>

namespace tnt_fuse {
namespace {

struct Axis {
  int64_t ext;    // extent (> 1)
  int64_t core;   // stride in the core matrix (rows: elements; columns: in units of m)
  int64_t other;  // stride in the payload on the other side of the transform
};

struct Piece {
  bool set = false;
  int64_t offset = 0;
  std::vector<int64_t> strides;  // one per row-box axis
  double coeff = 1.0;
};

struct Group {
  std::vector<std::pair<int64_t, int64_t>> box;  // row box: (extent, core stride), by core stride
  std::vector<Piece> src, dst;                   // k pieces, n pieces
};

struct Job {
  int64_t d_off, a_off, b_off, m, k, n, ldb;
};

int64_t count_of(const tnt_layout& l) {
  int64_t c = 1;
  for (int i = 0; i < l.rank; ++i) c *= l.shape[i];
  return c;
}

bool in_bounds(const tnt_layout& l, int64_t len) {
  if (l.rank < 0 || l.rank > TNT_MAX_RANK) return false;
  int64_t hi = l.offset;
  for (int i = 0; i < l.rank; ++i) {
    if (l.shape[i] < 0 || l.strides[i] < 0) return false;
    if (l.shape[i] > 0) hi += (l.shape[i] - 1) * l.strides[i];
  }
  return count_of(l) == 0 || (l.offset >= 0 && hi < len);
}

// The job whose matrix [start, start + rows * cols) holds `off` (starts sorted).
int find_job(const std::vector<std::pair<int64_t, int>>& starts, const std::vector<Job>& jobs, int64_t off,
             bool lhs_side) {
  auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(off, 1 << 30));
  if (it == starts.begin()) return -1;
  --it;
  const Job& j = jobs[it->second];
  const int64_t size = j.m * (lhs_side ? j.k : j.n);
  return off < it->first + size ? it->second : -1;
}

// Split a block of a job's column-major m-row matrix (layout `core`, matrix at
// `base`) into its row box and column box; `other` is the layout on the other
// side of the transform (same axes).
bool split_axes(const tnt_layout& core, const tnt_layout& other, int64_t base, int64_t m, int64_t ncols,
                int64_t& r0, int64_t& c0, std::vector<Axis>& rows, std::vector<Axis>& cols) {
  const int64_t p = core.offset - base;
  r0 = p % m;
  c0 = p / m;
  int64_t rhi = r0, chi = c0;
  for (int a = 0; a < core.rank; ++a) {
    const int64_t ext = core.shape[a], s = core.strides[a];
    if (ext == 1) continue;
    if (s < m) {
      rows.push_back({ext, s, other.strides[a]});
      rhi += (ext - 1) * s;
    } else {
      if (s % m) return false;
      cols.push_back({ext, s / m, other.strides[a]});
      chi += (ext - 1) * (s / m);
    }
  }
  if (rhi >= m || chi >= ncols) return false;
  std::stable_sort(rows.begin(), rows.end(), [](const Axis& x, const Axis& y) { return x.core < y.core; });
  return true;
}

std::vector<std::pair<int64_t, int64_t>> box_of(const std::vector<Axis>& rows) {
  std::vector<std::pair<int64_t, int64_t>> b;
  for (auto& r : rows) b.push_back({r.ext, r.core});
  return b;
}

// Visit every multi-index of the column box: f(column offset, other-side offset).
template <class F>
void for_each_col(const std::vector<Axis>& cols, F&& f) {
  std::vector<int64_t> idx(cols.size(), 0);
  while (true) {
    int64_t c = 0, o = 0;
    for (size_t i = 0; i < cols.size(); ++i) {
      c += idx[i] * cols[i].core;
      o += idx[i] * cols[i].other;
    }
    f(c, o);
    size_t i = 0;
    for (; i < cols.size(); ++i) {
      if (++idx[i] < cols[i].ext) break;
      idx[i] = 0;
    }
    if (i == cols.size()) return;
  }
}

bool single_only(const tnt_transform_desc* t) {
  for (size_t i = 0; i < t->n_blocks; ++i)
    if (t->blocks[i].kind != TNT_BLOCK_SINGLE || t->blocks[i].dst_count != 1 || t->blocks[i].src_count != 1 ||
        t->blocks[i].coeff >= t->n_coeffs || t->blocks[i].dst_layout >= t->n_layouts ||
        t->blocks[i].src_layout >= t->n_layouts)
      return false;
  return true;
}

bool fail(NarrowFusion& f, const char* why) {
  f.reason = why;
  return false;
}

}  // namespace

bool plan_narrow_fusion(const tnt_contract_desc* d, NarrowFusion& f) {
  f = NarrowFusion{};
  if (!d || d->n_jobs == 0) return fail(f, "no core jobs");
  if (d->lhs_op != TNT_OP_IDENTITY) return fail(f, "lhs op is not identity");
  if (!d->lhs && !d->out) return fail(f, "neither lhs nor output transform (plain GEMM)");
  if (d->lhs && (!single_only(d->lhs) || d->lhs->n_inactive))
    return fail(f, "lhs transform has Multi blocks or inactive layouts");
  if (d->rhs && !single_only(d->rhs)) return fail(f, "rhs transform has Multi blocks");
  if (d->out && (!single_only(d->out) || d->n_uncovered))
    return fail(f, "output transform has Multi blocks, or uncovered core regions");
  if (d->rhs_len > kMaxRhs) return fail(f, "rhs too large to map on the host");
  const bool rhs_adj = d->rhs_op == TNT_OP_TRANSPOSE || d->rhs_op == TNT_OP_ADJOINT;  // real payloads

  std::vector<Job> jobs;
  for (size_t i = 0; i < d->n_jobs; ++i) {
    const tnt_gemm_job& j = d->jobs[i];
    const int64_t m = (int64_t)j.rows, k = (int64_t)j.contracted, n = (int64_t)j.cols;
    if (m < 1 || k < 1 || n < 1 || k > kMaxK || n > kMaxN) return fail(f, "a job is not narrow (1 <= k <= 32, 1 <= n <= 16)");
    const int64_t ldb = rhs_adj ? n : k;
    if ((int64_t)j.lhs_offset + m * k > d->lhs_len || (int64_t)j.dst_offset + m * n > d->core_len ||
        (int64_t)j.rhs_offset + k * n > d->rhs_len)
      return fail(f, "job outside its operands");
    jobs.push_back({(int64_t)j.dst_offset, (int64_t)j.lhs_offset, (int64_t)j.rhs_offset, m, k, n, ldb});
  }
  std::vector<std::pair<int64_t, int>> a_starts, d_starts;
  for (size_t i = 0; i < jobs.size(); ++i) {
    a_starts.push_back({jobs[i].a_off, (int)i});
    d_starts.push_back({jobs[i].d_off, (int)i});
  }
  std::sort(a_starts.begin(), a_starts.end());
  std::sort(d_starts.begin(), d_starts.end());

  std::map<std::pair<int, int64_t>, Group> groups;  // (job, first row) -> group
  auto group_for = [&](int job, int64_t r0, const std::vector<Axis>& rows, bool create) -> Group* {
    auto key = std::make_pair(job, r0);
    auto it = groups.find(key);
    if (it == groups.end()) {
      if (!create) return nullptr;
      Group g;
      g.box = box_of(rows);
      g.src.resize((size_t)jobs[job].k);
      g.dst.resize((size_t)jobs[job].n);
      it = groups.emplace(key, std::move(g)).first;
    } else if (it->second.box != box_of(rows)) {
      return nullptr;
    }
    return &it->second;
  };

  if (d->lhs) {
    const tnt_transform_desc* t = d->lhs;
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      const tnt_layout& Lc = t->layouts[b.dst_layout];  // in the core lhs
      const tnt_layout& Ls = t->layouts[b.src_layout];  // in the lhs payload
      if (Lc.rank != Ls.rank || !in_bounds(Lc, d->lhs_len) || !in_bounds(Ls, t->src_len))
        return fail(f, "lhs layout out of bounds");
      for (int a = 0; a < Lc.rank; ++a)
        if (Lc.shape[a] != Ls.shape[a]) return fail(f, "lhs layouts differ in shape");
      if (count_of(Lc) == 0) continue;
      const int ji = find_job(a_starts, jobs, Lc.offset, true);
      if (ji < 0) return fail(f, "lhs block outside every job");
      const Job& J = jobs[ji];
      int64_t r0, c0;
      std::vector<Axis> rows, cols;
      if (!split_axes(Lc, Ls, J.a_off, J.m, J.k, r0, c0, rows, cols))
        return fail(f, "lhs block is not a row box x column box");
      Group* g = group_for(ji, r0, rows, true);
      if (!g) return fail(f, "lhs row trees disagree");
      const double c = t->coeffs[b.coeff];
      bool dup = false;
      for_each_col(cols, [&](int64_t dc, int64_t so) {
        Piece& p = g->src[(size_t)(c0 + dc)];
        if (p.set) dup = true;
        p.set = true;
        p.offset = Ls.offset + so;
        p.strides.clear();
        for (auto& r : rows) p.strides.push_back(r.other);
        p.coeff = c;
      });
      if (dup) return fail(f, "lhs blocks overlap");
    }
  }
  if (d->out) {
    const tnt_transform_desc* t = d->out;
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      const tnt_layout& Od = t->layouts[b.dst_layout];  // in the output payload
      const tnt_layout& Oc = t->layouts[b.src_layout];  // in the core destination
      if (Od.rank != Oc.rank || !in_bounds(Oc, d->core_len) || !in_bounds(Od, t->dst_len))
        return fail(f, "output layout out of bounds");
      for (int a = 0; a < Od.rank; ++a)
        if (Od.shape[a] != Oc.shape[a]) return fail(f, "output layouts differ in shape");
      if (count_of(Oc) == 0) continue;
      const int ji = find_job(d_starts, jobs, Oc.offset, false);
      if (ji < 0) return fail(f, "output block outside every job");
      const Job& J = jobs[ji];
      int64_t r0, c0;
      std::vector<Axis> rows, cols;
      if (!split_axes(Oc, Od, J.d_off, J.m, J.n, r0, c0, rows, cols))
        return fail(f, "output block is not a row box x column box");
      Group* g = group_for(ji, r0, rows, d->lhs == nullptr);
      if (!g) return fail(f, "output row trees disagree with the lhs row trees");
      const double c = t->coeffs[b.coeff];
      bool dup = false;
      for_each_col(cols, [&](int64_t dc, int64_t doff) {
        Piece& p = g->dst[(size_t)(c0 + dc)];
        if (p.set) dup = true;
        p.set = true;
        p.offset = Od.offset + doff;
        p.strides.clear();
        for (auto& r : rows) p.strides.push_back(r.other);
        p.coeff = c;
      });
      if (dup) return fail(f, "output blocks overlap");
    }
  }
  // the untransformed side addresses the core matrix itself (leading dimension m)
  std::vector<int64_t> rows_covered(jobs.size(), 0);
  for (auto& kv : groups) {
    const Job& J = jobs[kv.first.first];
    const int64_t r0 = kv.first.second;
    Group& g = kv.second;
    std::vector<int64_t> core_strides;
    int64_t cnt = 1;
    for (auto& e : g.box) {
      core_strides.push_back(e.second);
      cnt *= e.first;
    }
    rows_covered[kv.first.first] += cnt;
    if (!d->lhs)
      for (int64_t k = 0; k < J.k; ++k) g.src[(size_t)k] = {true, J.a_off + r0 + k * J.m, core_strides, 1.0};
    if (!d->out)
      for (int64_t n = 0; n < J.n; ++n) g.dst[(size_t)n] = {true, J.d_off + r0 + n * J.m, core_strides, 1.0};
    for (auto& p : g.src)
      if (!p.set) return fail(f, "a row tree misses a contracted column");
    for (auto& p : g.dst)
      if (!p.set) return fail(f, "a row tree misses an output column");
  }
  for (size_t i = 0; i < jobs.size(); ++i)
    if (rows_covered[i] != jobs[i].m) return fail(f, "row trees do not cover a job's rows");

  // core rhs element -> (rhs payload element, coefficient)
  std::vector<int64_t> ridx;
  std::vector<double> rco;
  const int64_t rhs_src_len = d->rhs ? d->rhs->src_len : d->rhs_len;
  if (d->rhs) {
    const tnt_transform_desc* t = d->rhs;
    ridx.assign((size_t)d->rhs_len, -1);
    rco.assign((size_t)d->rhs_len, 0.0);
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      const tnt_layout& Rc = t->layouts[b.dst_layout];
      const tnt_layout& Rs = t->layouts[b.src_layout];
      if (Rc.rank != Rs.rank || !in_bounds(Rc, d->rhs_len) || !in_bounds(Rs, t->src_len))
        return fail(f, "rhs layout out of bounds");
      std::vector<Axis> all;
      for (int a = 0; a < Rc.rank; ++a) {
        if (Rc.shape[a] != Rs.shape[a]) return fail(f, "rhs layouts differ in shape");
        all.push_back({Rc.shape[a], Rc.strides[a], Rs.strides[a]});
      }
      if (count_of(Rc) == 0) continue;
      const double c = t->coeffs[b.coeff];
      for_each_col(all, [&](int64_t q, int64_t s) {
        ridx[(size_t)(Rc.offset + q)] = Rs.offset + s;
        rco[(size_t)(Rc.offset + q)] = c;
      });
    }
  }

  for (auto& kv : groups) {
    const Job& J = jobs[kv.first.first];
    const Group& g = kv.second;
    tnt_transform_block blk{};
    blk.kind = TNT_BLOCK_MULTI;
    blk.dst_layout = (uint32_t)f.layouts.size();
    blk.dst_count = (uint32_t)J.n;
    blk.src_layout = (uint32_t)(f.layouts.size() + J.n);
    blk.src_count = (uint32_t)J.k;
    blk.coeff = (uint32_t)f.coeffs.size();
    auto push = [&](const Piece& p) {
      tnt_layout l{};
      l.offset = p.offset;
      l.rank = (int32_t)g.box.size();
      for (size_t i = 0; i < g.box.size(); ++i) {
        l.shape[i] = g.box[i].first;
        l.strides[i] = p.strides[i];
      }
      f.layouts.push_back(l);
    };
    for (auto& p : g.dst) push(p);
    for (auto& p : g.src) push(p);
    if (g.box.size() > (size_t)TNT_MAX_RANK) return fail(f, "row box rank exceeds the layout rank");
    for (int64_t n = 0; n < J.n; ++n)
      for (int64_t k = 0; k < J.k; ++k) {
        const int64_t q = J.b_off + (rhs_adj ? n + k * J.ldb : k + n * J.ldb);
        int64_t idx = d->rhs ? ridx[(size_t)q] : q;
        double s = g.dst[(size_t)n].coeff * g.src[(size_t)k].coeff * (d->rhs ? rco[(size_t)q] : 1.0);
        if (idx < 0) s = 0.0;
        f.gather_idx.push_back(idx);
        f.gather_scale.push_back(s);
        f.coeffs.push_back(0.0);
      }
    f.blocks.push_back(blk);
  }
  // output elements written by no group: inactive output layouts, or core
  // regions written by no job (only without an output transform)
  if (d->out) {
    for (size_t i = 0; i < d->out->n_inactive; ++i) {
      if (d->out->inactive[i] >= d->out->n_layouts) return fail(f, "inactive layout index out of range");
      f.inactive.push_back((uint32_t)f.layouts.size());
      f.layouts.push_back(d->out->layouts[d->out->inactive[i]]);
    }
  } else {
    for (size_t i = 0; i < d->n_uncovered; ++i) {
      tnt_layout l{};
      l.offset = d->uncovered[i].offset;
      l.rank = 1;
      l.shape[0] = d->uncovered[i].len;
      l.strides[0] = 1;
      f.inactive.push_back((uint32_t)f.layouts.size());
      f.layouts.push_back(l);
    }
  }
  f.rhs_src_len = rhs_src_len;
  f.desc.dst_len = d->out ? d->out->dst_len : d->core_len;
  f.desc.src_len = d->lhs ? d->lhs->src_len : d->lhs_len;
  f.desc.n_layouts = f.layouts.size();
  f.desc.layouts = f.layouts.data();
  f.desc.n_blocks = f.blocks.size();
  f.desc.blocks = f.blocks.data();
  f.desc.n_coeffs = f.coeffs.size();
  f.desc.coeffs = f.coeffs.data();
  f.desc.n_inactive = f.inactive.size();
  f.desc.inactive = f.inactive.data();
  return true;
}

}  // namespace tnt_fuse
