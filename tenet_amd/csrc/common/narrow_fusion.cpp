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

bool blocks_valid(const tnt_transform_desc* t) {
  for (size_t i = 0; i < t->n_blocks; ++i) {
    const tnt_transform_block& b = t->blocks[i];
    if (b.kind != TNT_BLOCK_SINGLE && b.kind != TNT_BLOCK_MULTI) return false;
    if (b.kind == TNT_BLOCK_SINGLE && (b.dst_count != 1 || b.src_count != 1)) return false;
    if (b.dst_count == 0 || b.src_count == 0) return false;
    const uint64_t nc = b.kind == TNT_BLOCK_SINGLE ? 1 : (uint64_t)b.dst_count * b.src_count;
    if ((uint64_t)b.coeff + nc > t->n_coeffs || (uint64_t)b.dst_layout + b.dst_count > t->n_layouts ||
        (uint64_t)b.src_layout + b.src_count > t->n_layouts)
      return false;
  }
  return true;
}

// U[dst j, src i] of a block (Single: its one coefficient).
double block_coeff(const tnt_transform_desc* t, const tnt_transform_block& b, uint32_t j, uint32_t i) {
  return b.kind == TNT_BLOCK_SINGLE ? t->coeffs[b.coeff] : t->coeffs[b.coeff + (size_t)j * b.src_count + i];
}

bool fail(NarrowFusion& f, const char* why) {
  f.reason = why;
  return false;
}

// A piece of a payload shaped like a row box: offset and one stride per box axis.
struct PieceKey {
  int64_t offset;
  std::vector<int64_t> strides;
  bool operator<(const PieceKey& o) const { return offset != o.offset ? offset < o.offset : strides < o.strides; }
};

struct LhsTerm {  // coefficient x source piece
  double coef;
  PieceKey piece;
};

struct OutTerm {  // coefficient x core column slice (node, column)
  double coef;
  int node;
  int64_t col;
};

struct OutPiece {
  PieceKey piece;
  std::vector<OutTerm> terms;
};

struct Node {
  int job;
  int64_t r0;
  std::vector<std::pair<int64_t, int64_t>> box;   // (extent, core stride), by core stride
  std::vector<std::vector<LhsTerm>> cols;         // per contracted column: its terms
  std::vector<bool> col_set;
};

struct UnionFind {
  std::vector<int> p;
  int add() {
    p.push_back((int)p.size());
    return (int)p.size() - 1;
  }
  int find(int x) {
    while (p[x] != x) x = p[x] = p[p[x]];
    return x;
  }
  void join(int a, int b) {
    a = find(a);
    b = find(b);
    if (a != b) p[std::max(a, b)] = std::min(a, b);
  }
};

}  // namespace

// Nodes are (job, row box) pieces of the core: the rows of one row tree of a
// coupled sector.  Single blocks keep a node on its own; a Multi block mixes
// the nodes of its core-side layouts (lhs: its destinations, output: its
// sources), so the fused groups are the connected components of nodes under
// the Multi blocks of both transforms.  Per component, every output piece is
//   out_l = sum_s U[l][s] src_s,
//   U[l][s] = sum over the output terms (c_o, node, n), the contracted
//             columns k of the node's job, the rhs terms (c_r, rhs element)
//             of core rhs element (k, n) and the lhs terms (c_l, src_s) of
//             the node's column k, of c_o c_r c_l rhs[element],
// a sum of products of plan coefficients and rhs payload elements, gathered
// on the device before the pass.  Without Multi blocks every component is
// one node and every U entry one term: the groups of the Single-only plan.
bool plan_narrow_fusion(const tnt_contract_desc* d, NarrowFusion& f) {
  f = NarrowFusion{};
  if (!d || d->n_jobs == 0) return fail(f, "no core jobs");
  if (d->lhs_op != TNT_OP_IDENTITY) return fail(f, "lhs op is not identity");
  if (!d->lhs && !d->out) return fail(f, "neither lhs nor output transform (plain GEMM)");
  if (d->lhs && (!blocks_valid(d->lhs) || d->lhs->n_inactive))
    return fail(f, "lhs transform has malformed blocks or inactive layouts");
  if (d->rhs && !blocks_valid(d->rhs)) return fail(f, "rhs transform has malformed blocks");
  if (d->out && (!blocks_valid(d->out) || d->n_uncovered))
    return fail(f, "output transform has malformed blocks, or uncovered core regions");
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

  std::vector<Node> nodes;
  std::map<std::pair<int, int64_t>, int> node_of;  // (job, first row) -> node
  UnionFind uf;
  // the node of a core block; created only on the side that defines the nodes
  auto node_for = [&](int job, int64_t r0, const std::vector<Axis>& rows, bool create) -> int {
    auto key = std::make_pair(job, r0);
    auto it = node_of.find(key);
    if (it == node_of.end()) {
      if (!create) return -1;
      Node nd;
      nd.job = job;
      nd.r0 = r0;
      nd.box = box_of(rows);
      nd.cols.resize((size_t)jobs[job].k);
      nd.col_set.assign((size_t)jobs[job].k, false);
      nodes.push_back(std::move(nd));
      uf.add();
      it = node_of.emplace(key, (int)nodes.size() - 1).first;
    } else if (nodes[it->second].box != box_of(rows)) {
      return -2;
    }
    return it->second;
  };
  // the other-side layout of a block, with its axes split like the core layout's
  struct Split {
    int job = -1;
    int64_t r0 = 0, c0 = 0;
    std::vector<Axis> rows, cols;
  };
  auto split_with = [&](const tnt_layout& core, const tnt_layout& other, bool lhs_side, Split& sp) -> bool {
    const auto& starts = lhs_side ? a_starts : d_starts;
    sp.job = find_job(starts, jobs, core.offset, lhs_side);
    if (sp.job < 0) return false;
    const Job& J = jobs[sp.job];
    sp.rows.clear();
    sp.cols.clear();
    return split_axes(core, other, lhs_side ? J.a_off : J.d_off, J.m, lhs_side ? J.k : J.n, sp.r0, sp.c0, sp.rows,
                      sp.cols);
  };

  if (d->lhs) {
    const tnt_transform_desc* t = d->lhs;
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      // every layout of a block has the same shape; each destination is a core block
      const tnt_layout& L0 = t->layouts[b.dst_layout];
      for (uint32_t q = 0; q < b.dst_count + b.src_count; ++q) {
        const bool dst = q < b.dst_count;
        const tnt_layout& L = t->layouts[dst ? b.dst_layout + q : b.src_layout + (q - b.dst_count)];
        if (L.rank != L0.rank || !in_bounds(L, dst ? d->lhs_len : t->src_len))
          return fail(f, "lhs layout out of bounds");
        for (int a = 0; a < L.rank; ++a)
          if (L.shape[a] != L0.shape[a]) return fail(f, "lhs layouts differ in shape");
      }
      if (count_of(L0) == 0) continue;
      int first_node = -1;
      for (uint32_t jd = 0; jd < b.dst_count; ++jd) {
        const tnt_layout& Lc = t->layouts[b.dst_layout + jd];
        Split sp;  // split the core layout; the source strides are read per source below
        if (!split_with(Lc, Lc, true, sp)) return fail(f, "lhs block is not a row box x column box of a job");
        const int nd = node_for(sp.job, sp.r0, sp.rows, true);
        if (nd < 0) return fail(f, "lhs row trees disagree");
        if (first_node < 0) first_node = nd;
        else uf.join(first_node, nd);
        // axes of the core layout in the order split_axes found them
        std::vector<int> row_axes, col_axes;
        for (int a = 0; a < Lc.rank; ++a) {
          if (Lc.shape[a] == 1) continue;
          (Lc.strides[a] < jobs[sp.job].m ? row_axes : col_axes).push_back(a);
        }
        std::stable_sort(row_axes.begin(), row_axes.end(),
                         [&](int x, int y) { return Lc.strides[x] < Lc.strides[y]; });
        Node& N = nodes[nd];
        bool dup = false;
        // visit the column box: core column offset and, per source, the source offset
        std::vector<int64_t> idx(col_axes.size(), 0);
        while (true) {
          int64_t dc = 0;
          for (size_t u = 0; u < col_axes.size(); ++u) dc += idx[u] * (Lc.strides[col_axes[u]] / jobs[sp.job].m);
          const int64_t c = sp.c0 + dc;
          if (N.col_set[(size_t)c]) dup = true;
          N.col_set[(size_t)c] = true;
          for (uint32_t is = 0; is < b.src_count; ++is) {
            const tnt_layout& Ls = t->layouts[b.src_layout + is];
            PieceKey pk;
            pk.offset = Ls.offset;
            for (size_t u = 0; u < col_axes.size(); ++u) pk.offset += idx[u] * Ls.strides[col_axes[u]];
            for (int a : row_axes) pk.strides.push_back(Ls.strides[a]);
            N.cols[(size_t)c].push_back({block_coeff(t, b, jd, is), pk});
          }
          size_t u = 0;
          for (; u < col_axes.size(); ++u) {
            if (++idx[u] < Lc.shape[col_axes[u]]) break;
            idx[u] = 0;
          }
          if (u == col_axes.size()) break;
        }
        if (dup) return fail(f, "lhs blocks overlap");
      }
    }
  }

  std::vector<OutPiece> outs;
  if (d->out) {
    const tnt_transform_desc* t = d->out;
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      const tnt_layout& L0 = t->layouts[b.dst_layout];
      for (uint32_t q = 0; q < b.dst_count + b.src_count; ++q) {
        const bool dst = q < b.dst_count;
        const tnt_layout& L = t->layouts[dst ? b.dst_layout + q : b.src_layout + (q - b.dst_count)];
        if (L.rank != L0.rank || !in_bounds(L, dst ? t->dst_len : d->core_len))
          return fail(f, "output layout out of bounds");
        for (int a = 0; a < L.rank; ++a)
          if (L.shape[a] != L0.shape[a]) return fail(f, "output layouts differ in shape");
      }
      if (count_of(L0) == 0) continue;
      // the core-side (source) layouts: their nodes and axis splits
      std::vector<int> src_node(b.src_count);
      std::vector<std::vector<int>> row_axes(b.src_count), col_axes(b.src_count);
      std::vector<int64_t> c0s(b.src_count);
      for (uint32_t is = 0; is < b.src_count; ++is) {
        const tnt_layout& Oc = t->layouts[b.src_layout + is];
        Split sp;
        if (!split_with(Oc, Oc, false, sp)) return fail(f, "output block is not a row box x column box of a job");
        const int nd = node_for(sp.job, sp.r0, sp.rows, d->lhs == nullptr);
        if (nd == -2) return fail(f, "output row trees disagree with the lhs row trees");
        if (nd < 0) return fail(f, "an output block reads rows no lhs block writes");
        src_node[is] = nd;
        c0s[is] = sp.c0;
        if (is > 0) uf.join(src_node[0], nd);
        for (int a = 0; a < Oc.rank; ++a) {
          if (Oc.shape[a] == 1) continue;
          (Oc.strides[a] < jobs[sp.job].m ? row_axes[is] : col_axes[is]).push_back(a);
        }
        std::stable_sort(row_axes[is].begin(), row_axes[is].end(),
                         [&](int x, int y) { return Oc.strides[x] < Oc.strides[y]; });
        if (row_axes[is] != row_axes[0] || col_axes[is] != col_axes[0])
          return fail(f, "output sources split their axes differently");
      }
      const std::vector<int>& ra = row_axes[0];
      const std::vector<int>& ca = col_axes[0];
      std::vector<int64_t> idx(ca.size(), 0);
      while (true) {
        for (uint32_t jd = 0; jd < b.dst_count; ++jd) {
          const tnt_layout& Od = t->layouts[b.dst_layout + jd];
          OutPiece op;
          op.piece.offset = Od.offset;
          for (size_t u = 0; u < ca.size(); ++u) op.piece.offset += idx[u] * Od.strides[ca[u]];
          for (int a : ra) op.piece.strides.push_back(Od.strides[a]);
          for (uint32_t is = 0; is < b.src_count; ++is) {
            const tnt_layout& Oc = t->layouts[b.src_layout + is];
            const int64_t m = jobs[nodes[src_node[is]].job].m;
            int64_t dc = 0;
            for (size_t u = 0; u < ca.size(); ++u) dc += idx[u] * (Oc.strides[ca[u]] / m);
            op.terms.push_back({block_coeff(t, b, jd, is), src_node[is], c0s[is] + dc});
          }
          outs.push_back(std::move(op));
        }
        size_t u = 0;
        for (; u < ca.size(); ++u) {
          if (++idx[u] < L0.shape[ca[u]]) break;
          idx[u] = 0;
        }
        if (u == ca.size()) break;
      }
    }
  }

  // the untransformed side addresses the core matrix itself (leading dimension m)
  std::vector<int64_t> rows_covered(jobs.size(), 0);
  for (size_t ni = 0; ni < nodes.size(); ++ni) {
    Node& N = nodes[ni];
    const Job& J = jobs[N.job];
    std::vector<int64_t> core_strides;
    int64_t cnt = 1;
    for (auto& e : N.box) {
      core_strides.push_back(e.second);
      cnt *= e.first;
    }
    rows_covered[N.job] += cnt;
    if (!d->lhs)
      for (int64_t k = 0; k < J.k; ++k) {
        N.cols[(size_t)k] = {{1.0, PieceKey{J.a_off + N.r0 + k * J.m, core_strides}}};
        N.col_set[(size_t)k] = true;
      }
    if (!d->out)
      for (int64_t n = 0; n < J.n; ++n)
        outs.push_back({PieceKey{J.d_off + N.r0 + n * J.m, core_strides}, {{1.0, (int)ni, n}}});
    for (bool s : N.col_set)
      if (!s) return fail(f, "a row tree misses a contracted column");
  }
  for (size_t i = 0; i < jobs.size(); ++i)
    if (rows_covered[i] != jobs[i].m) return fail(f, "row trees do not cover a job's rows");

  // core rhs element -> terms (rhs payload element, coefficient), sorted by element
  struct RhsTerm {
    int64_t q, idx;
    double c;
  };
  std::vector<RhsTerm> rterms;
  const int64_t rhs_src_len = d->rhs ? d->rhs->src_len : d->rhs_len;
  if (d->rhs) {
    const tnt_transform_desc* t = d->rhs;
    for (size_t bi = 0; bi < t->n_blocks; ++bi) {
      const tnt_transform_block& b = t->blocks[bi];
      const tnt_layout& R0 = t->layouts[b.dst_layout];
      for (uint32_t q = 0; q < b.dst_count + b.src_count; ++q) {
        const bool dst = q < b.dst_count;
        const tnt_layout& R = t->layouts[dst ? b.dst_layout + q : b.src_layout + (q - b.dst_count)];
        if (R.rank != R0.rank || !in_bounds(R, dst ? d->rhs_len : t->src_len))
          return fail(f, "rhs layout out of bounds");
        for (int a = 0; a < R.rank; ++a)
          if (R.shape[a] != R0.shape[a]) return fail(f, "rhs layouts differ in shape");
      }
      if (count_of(R0) == 0) continue;
      for (uint32_t jd = 0; jd < b.dst_count; ++jd)
        for (uint32_t is = 0; is < b.src_count; ++is) {
          const tnt_layout& Rc = t->layouts[b.dst_layout + jd];
          const tnt_layout& Rs = t->layouts[b.src_layout + is];
          std::vector<Axis> all;
          for (int a = 0; a < Rc.rank; ++a) all.push_back({Rc.shape[a], Rc.strides[a], Rs.strides[a]});
          const double c = block_coeff(t, b, jd, is);
          for_each_col(all, [&](int64_t qo, int64_t so) { rterms.push_back({Rc.offset + qo, Rs.offset + so, c}); });
        }
    }
    std::stable_sort(rterms.begin(), rterms.end(), [](const RhsTerm& x, const RhsTerm& y) { return x.q < y.q; });
  }
  auto rhs_terms = [&](int64_t q, std::vector<std::pair<int64_t, double>>& out) {
    out.clear();
    if (!d->rhs) {
      out.push_back({q, 1.0});
      return;
    }
    auto it = std::lower_bound(rterms.begin(), rterms.end(), q, [](const RhsTerm& x, int64_t v) { return x.q < v; });
    for (; it != rterms.end() && it->q == q; ++it) out.push_back({it->idx, it->c});
  };

  // components: nodes in (job, first row) order; output pieces by their first term's node
  std::map<int, std::vector<int>> comp_nodes;
  for (auto& kv : node_of) comp_nodes[uf.find(kv.second)].push_back(kv.second);
  std::map<int, std::vector<size_t>> comp_outs;
  for (size_t l = 0; l < outs.size(); ++l) {
    if (outs[l].terms.empty()) continue;
    comp_outs[uf.find(outs[l].terms[0].node)].push_back(l);
  }
  std::vector<std::pair<int64_t, double>> rt;
  f.gather_ptr.push_back(0);
  for (auto& kv : comp_nodes) {
    const std::vector<int>& cn = kv.second;
    const auto& box = nodes[cn[0]].box;
    for (int nd : cn)
      for (size_t a = 0; a < box.size(); ++a)
        if (nodes[nd].box.size() != box.size() || nodes[nd].box[a].first != box[a].first)
          return fail(f, "a fused group mixes row trees of different shapes");
    if (box.size() > (size_t)TNT_MAX_RANK) return fail(f, "row box rank exceeds the layout rank");
    // sources: the lhs pieces of every column of every node, first use first
    std::vector<PieceKey> srcs;
    std::map<PieceKey, int> src_id;
    for (int nd : cn)
      for (auto& col : nodes[nd].cols)
        for (auto& tm : col)
          if (src_id.emplace(tm.piece, (int)srcs.size()).second) srcs.push_back(tm.piece);
    const std::vector<size_t>& ol = comp_outs[kv.first];
    if (ol.empty()) continue;  // core rows nothing reads
    if ((int64_t)srcs.size() > kMaxK) return fail(f, "a fused group has more than 32 sources");
    tnt_transform_block blk{};
    blk.kind = TNT_BLOCK_MULTI;
    blk.dst_layout = (uint32_t)f.layouts.size();
    blk.dst_count = (uint32_t)ol.size();
    blk.src_layout = (uint32_t)(f.layouts.size() + ol.size());
    blk.src_count = (uint32_t)srcs.size();
    blk.coeff = (uint32_t)f.coeffs.size();
    auto push = [&](const PieceKey& p) {
      tnt_layout l{};
      l.offset = p.offset;
      l.rank = (int32_t)box.size();
      for (size_t i = 0; i < box.size(); ++i) {
        l.shape[i] = box[i].first;
        l.strides[i] = p.strides[i];
      }
      f.layouts.push_back(l);
    };
    for (size_t l : ol) push(outs[l].piece);
    for (auto& p : srcs) push(p);
    for (size_t l : ol) {
      // U[l][s] as (rhs element, scale) terms, accumulated per source
      std::vector<std::vector<std::pair<int64_t, double>>> u(srcs.size());
      for (const OutTerm& ot : outs[l].terms) {
        const Node& N = nodes[ot.node];
        const Job& J = jobs[N.job];
        if (ot.col < 0 || ot.col >= J.n) return fail(f, "output term outside its job");
        for (int64_t k = 0; k < J.k; ++k) {
          const int64_t q = J.b_off + (rhs_adj ? ot.col + k * J.ldb : k + ot.col * J.ldb);
          rhs_terms(q, rt);
          for (const LhsTerm& lt : N.cols[(size_t)k]) {
            auto& cell = u[(size_t)src_id.at(lt.piece)];
            for (auto& r : rt) {
              const double sc = ot.coef * lt.coef * r.second;  // c_out c_lhs c_rhs
              if (r.first < 0 || sc == 0.0) continue;
              bool merged = false;
              for (auto& e : cell)
                if (e.first == r.first) {
                  e.second += sc;
                  merged = true;
                  break;
                }
              if (!merged) cell.push_back({r.first, sc});
            }
          }
        }
      }
      for (auto& cell : u) {
        for (auto& e : cell) {
          f.gather_idx.push_back(e.first);
          f.gather_scale.push_back(e.second);
        }
        f.gather_ptr.push_back((int64_t)f.gather_idx.size());
        f.coeffs.push_back(0.0);
      }
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
