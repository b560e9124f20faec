// Host structure compiler (cold path).  See structure.hpp.
//
// Reference correspondence (paths relative to the TeNeT repository root):
//   sector ids / symbols   tenet-sectors/src/abelian.rs:517-780, su2.rs:60-370,
//                          algebra.rs:552-625, :1434-1479 (Cantor product codec)
//   tree enumeration       tenet-core/src/fusion_tree.rs:491-575, :970-990, :1076-1131
//   block layout           tenet-core/src/fusion_space.rs:2211-2261, :3953-4104
//   Artin braid            tenet-core/src/fusion_tree.rs:2477-2505, :8088-8273
//   bends / repartition    tenet-core/src/fusion_tree.rs:3785-3812, :8749-9040
//   permute                tenet-core/src/fusion_tree.rs:3010-3120, :3278-3420, :7097-7130
//   group planning         tenet-tensors/src/tree_transform/plan.rs:1489-1540, :2107-2200
//   structure compile      tenet-operations/src/transform_structure.rs:587-776
//   contraction route      tenet-tensors/src/contract/dynamic.rs:1241-1530,
//                          tenet-operations/src/fusion_replay.rs:457-555
#include "structure.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <set>

namespace tnth {

// ---------------------------------------------------------------------------
// Sector ids
// ---------------------------------------------------------------------------
Sid u1_id(int64_t q) { return q >= 0 ? Sid(2 * q) : Sid(-2 * q - 1); }
int64_t u1_charge(Sid z) { return (z & 1) ? -int64_t((z + 1) / 2) : int64_t(z / 2); }

Sid cantor(Sid l, Sid r) {
  Sid s = l + r;
  return s * (s + 1) / 2 + l;
}

void uncantor(Sid id, Sid& l, Sid& r) {
  Sid s = Sid((std::sqrt(8.0L * (long double)id + 1.0L) - 1.0L) / 2.0L);
  while (s * (s + 1) / 2 > id) --s;
  while ((s + 1) * (s + 2) / 2 <= id) ++s;
  l = id - s * (s + 1) / 2;
  r = s - l;
}

// ---------------------------------------------------------------------------
// SU(2) 6j symbol (Racah formula) in long double
// ---------------------------------------------------------------------------
static long double lfact(int64_t n) {
  static std::vector<long double> table = [] {
    std::vector<long double> t(400);
    t[0] = 1.0L;
    for (size_t i = 1; i < t.size(); ++i) t[i] = t[i - 1] * (long double)i;
    return t;
  }();
  if (n < 0 || n >= (int64_t)table.size()) throw Error("factorial argument out of range");
  return table[n];
}

static bool triangle(int64_t a, int64_t b, int64_t c) {
  return ((a + b + c) % 2 == 0) && std::llabs(a - b) <= c && c <= a + b;
}

static long double delta2(int64_t a, int64_t b, int64_t c) {
  return lfact((a + b - c) / 2) * lfact((a - b + c) / 2) * lfact((-a + b + c) / 2) /
         lfact((a + b + c) / 2 + 1);
}

double sixj_doubled(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6) {
  if (!(triangle(j1, j2, j3) && triangle(j1, j5, j6) && triangle(j4, j2, j6) &&
        triangle(j4, j5, j3)))
    return 0.0;
  int64_t a1 = (j1 + j2 + j3) / 2, a2 = (j1 + j5 + j6) / 2, a3 = (j4 + j2 + j6) / 2,
          a4 = (j4 + j5 + j3) / 2;
  int64_t b1 = (j1 + j2 + j4 + j5) / 2, b2 = (j2 + j3 + j5 + j6) / 2,
          b3 = (j3 + j1 + j6 + j4) / 2;
  int64_t tmin = std::max(std::max(a1, a2), std::max(a3, a4));
  int64_t tmax = std::min(b1, std::min(b2, b3));
  // Sum positive and negative terms separately to limit cancellation error.
  long double pos = 0.0L, neg = 0.0L;
  for (int64_t t = tmin; t <= tmax; ++t) {
    long double term = lfact(t + 1) /
                       (lfact(t - a1) * lfact(t - a2) * lfact(t - a3) * lfact(t - a4) *
                        lfact(b1 - t) * lfact(b2 - t) * lfact(b3 - t));
    if (t % 2) neg += term; else pos += term;
  }
  long double total = pos - neg;
  long double d = delta2(j1, j2, j3) * delta2(j1, j5, j6) * delta2(j4, j2, j6) * delta2(j4, j5, j3);
  long double v = total * sqrtl(d);
  // Exact zeros of the alternating sum come out as rounding noise.
  if (fabsl(total) <= 1e-15L * (pos + neg)) return 0.0;
  return (double)v;
}

// ---------------------------------------------------------------------------
// Rule
// ---------------------------------------------------------------------------
Sid Rule::vacuum() const {
  switch (kind) {
    case RuleKind::U1xZ2: return cantor(0, 0);
    default: return 0;
  }
}

Sid Rule::dual(Sid a) const {
  switch (kind) {
    case RuleKind::U1: return u1_id(-u1_charge(a));
    case RuleKind::SU2: return a;
    case RuleKind::Z2: return a;
    case RuleKind::U1xZ2: {
      Sid l, r;
      uncantor(a, l, r);
      return cantor(u1_id(-u1_charge(l)), r);
    }
  }
  return a;
}

void Rule::channels(Sid a, Sid b, std::vector<Sid>& out) const {
  out.clear();
  switch (kind) {
    case RuleKind::U1: out.push_back(u1_id(u1_charge(a) + u1_charge(b))); break;
    case RuleKind::SU2: {
      int64_t lo = std::llabs((int64_t)a - (int64_t)b), hi = (int64_t)(a + b);
      for (int64_t c = lo; c <= hi; c += 2) out.push_back(Sid(c));
      break;
    }
    case RuleKind::Z2: out.push_back(a ^ b); break;
    case RuleKind::U1xZ2: {
      Sid al, ar, bl, br;
      uncantor(a, al, ar);
      uncantor(b, bl, br);
      out.push_back(cantor(u1_id(u1_charge(al) + u1_charge(bl)), ar ^ br));
      break;
    }
  }
}

int Rule::nsym(Sid a, Sid b, Sid c) const {
  switch (kind) {
    case RuleKind::U1: return u1_charge(a) + u1_charge(b) == u1_charge(c);
    case RuleKind::SU2: return triangle((int64_t)a, (int64_t)b, (int64_t)c);
    case RuleKind::Z2: return (a ^ b) == c;
    case RuleKind::U1xZ2: {
      Sid al, ar, bl, br, cl, cr;
      uncantor(a, al, ar);
      uncantor(b, bl, br);
      uncantor(c, cl, cr);
      return (u1_charge(al) + u1_charge(bl) == u1_charge(cl)) && ((ar ^ br) == cr);
    }
  }
  return 0;
}

double Rule::F(Sid a, Sid b, Sid c, Sid d, Sid e, Sid f) const {
  if (kind != RuleKind::SU2) return 1.0;
  int64_t A = a, B = b, C = c, D = d, E = e, Fv = f;
  if (!(triangle(A, B, E) && triangle(E, C, D) && triangle(B, C, Fv) && triangle(A, Fv, D))) return 0.0;
  if (A == 0 && B == 0 && C == 0 && D == 0 && E == 0 && Fv == 0) return 1.0;
  double sign = (((A + B + C + D) / 2) % 2) ? -1.0 : 1.0;
  return std::sqrt(double((E + 1) * (Fv + 1))) * sign * sixj_doubled(A, B, E, C, D, Fv);
}

double Rule::R(Sid a, Sid b, Sid c) const {
  if (kind != RuleKind::SU2) return 1.0;
  int64_t A = a, B = b, C = c;
  if (!triangle(A, B, C)) return 0.0;
  return (((A + B - C) / 2) % 2) ? -1.0 : 1.0;
}

double Rule::dim(Sid a) const { return kind == RuleKind::SU2 ? double(a + 1) : 1.0; }
double Rule::sqrtdim(Sid a) const { return std::sqrt(dim(a)); }
double Rule::fs(Sid a) const { return (kind == RuleKind::SU2 && (a % 2)) ? -1.0 : 1.0; }

double Rule::bsym(Sid a, Sid b, Sid c) const {
  return sqrtdim(a) * sqrtdim(b) / sqrtdim(c) * F(a, b, dual(b), a, c, vacuum());
}

// ---------------------------------------------------------------------------
// Legs, trees, hashing
// ---------------------------------------------------------------------------
int64_t Leg::deg(Sid s) const {
  auto it = std::lower_bound(sectors.begin(), sectors.end(), s);
  if (it == sectors.end() || *it != s) throw Error("sector absent from leg");
  return degs[it - sectors.begin()];
}

Leg Leg::dualized(const Rule& rule) const {
  std::vector<std::pair<Sid, int64_t>> p;
  for (size_t i = 0; i < sectors.size(); ++i) p.emplace_back(rule.dual(sectors[i]), degs[i]);
  std::sort(p.begin(), p.end());
  Leg out;
  for (auto& x : p) {
    out.sectors.push_back(x.first);
    out.degs.push_back(x.second);
  }
  out.is_dual = !is_dual;
  return out;
}

static inline void hmix(size_t& h, uint64_t v) {
  h ^= std::hash<uint64_t>()(v) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
}

static size_t tree_hash(const Tree& t) {
  size_t h = t.unc.size();
  for (auto s : t.unc) hmix(h, s);
  for (auto s : t.inner) hmix(h, s + 7919);
  for (auto d : t.isdual) hmix(h, d + 31);
  hmix(h, t.coupled + 104729);
  return h;
}

size_t PairKeyHash::operator()(const PairKey& k) const {
  size_t h = tree_hash(k.cod);
  hmix(h, tree_hash(k.dom));
  return h;
}

// ---------------------------------------------------------------------------
// Tree enumeration (fusion_tree.rs:491-575, 970-990, 1076-1131)
// ---------------------------------------------------------------------------
static void visit_inner(const Rule& rule, const std::vector<Sid>& eff, size_t n, Sid coupled,
                        std::vector<Sid>& stack, std::vector<std::vector<Sid>>& out) {
  // stack holds inner lines outermost-first; emitted lists are innermost-first.
  if (n == 0) {
    if (coupled == rule.vacuum()) out.emplace_back(stack.rbegin(), stack.rend());
    return;
  }
  if (n == 1) {
    if (eff[0] == coupled) out.emplace_back(stack.rbegin(), stack.rend());
    return;
  }
  if (n == 2) {
    if (rule.nsym(eff[0], eff[1], coupled)) out.emplace_back(stack.rbegin(), stack.rend());
    return;
  }
  Sid last = eff[n - 1];
  std::vector<Sid> ch;
  rule.channels(coupled, rule.dual(last), ch);
  for (Sid a : ch) {
    if (!rule.nsym(a, last, coupled)) continue;
    stack.push_back(a);
    visit_inner(rule, eff, n - 1, a, stack, out);
    stack.pop_back();
  }
}

static std::vector<Sid> reachable(const Rule& rule, const std::vector<Sid>& eff) {
  if (eff.empty()) return {rule.vacuum()};
  std::vector<Sid> acc{eff[0]}, ch;
  for (size_t i = 1; i < eff.size(); ++i) {
    std::set<Sid> next;
    for (Sid f : acc) {
      rule.channels(f, eff[i], ch);
      next.insert(ch.begin(), ch.end());
    }
    acc.assign(next.begin(), next.end());
  }
  return acc;
}

static std::map<Sid, std::vector<Tree>> trees_by_coupled(const Rule& rule,
                                                         const std::vector<Leg>& legs) {
  std::map<Sid, std::vector<Tree>> groups;
  size_t n = legs.size();
  std::vector<uint8_t> isd(n);
  for (size_t i = 0; i < n; ++i) isd[i] = legs[i].is_dual;
  for (auto& l : legs)
    if (l.sectors.empty()) return groups;
  std::vector<size_t> idx(n, 0);
  std::vector<Sid> tup(n), stack;
  std::vector<std::vector<Sid>> inners;
  while (true) {
    for (size_t i = 0; i < n; ++i) tup[i] = legs[i].sectors[idx[i]];
    for (Sid c : reachable(rule, tup)) {
      inners.clear();
      stack.clear();
      visit_inner(rule, tup, n, c, stack, inners);
      auto& g = groups[c];
      for (auto& in : inners) {
        Tree t;
        t.unc = tup;
        t.isdual = isd;
        t.coupled = c;
        t.inner = (n > 2) ? in : std::vector<Sid>();
        g.push_back(std::move(t));
      }
    }
    // first leg fastest
    size_t k = 0;
    while (k < n) {
      if (++idx[k] < legs[k].sectors.size()) break;
      idx[k] = 0;
      ++k;
    }
    if (k == n) break;
  }
  for (auto it = groups.begin(); it != groups.end();) {
    if (it->second.empty()) it = groups.erase(it); else ++it;
  }
  return groups;
}

void HomSpace::build() {
  blocks.clear();
  index.clear();
  sectors.clear();
  sector_index.clear();
  auto cg = trees_by_coupled(rule, cod);
  auto dg = trees_by_coupled(rule, dom);
  int64_t offset = 0;
  for (auto& kv : cg) {
    auto it = dg.find(kv.first);
    if (it == dg.end()) continue;
    const auto& ctrees = kv.second;
    const auto& dtrees = it->second;
    std::vector<std::vector<int64_t>> rsh(ctrees.size()), csh(dtrees.size());
    std::vector<int64_t> rstart(ctrees.size()), cstart(dtrees.size());
    int64_t rows = 0, cols = 0;
    for (size_t i = 0; i < ctrees.size(); ++i) {
      int64_t p = 1;
      for (size_t a = 0; a < cod.size(); ++a) {
        rsh[i].push_back(cod[a].deg(ctrees[i].unc[a]));
        p *= rsh[i].back();
      }
      rstart[i] = rows;
      rows += p;
    }
    for (size_t j = 0; j < dtrees.size(); ++j) {
      int64_t p = 1;
      for (size_t a = 0; a < dom.size(); ++a) {
        csh[j].push_back(dom[a].deg(dtrees[j].unc[a]));
        p *= csh[j].back();
      }
      cstart[j] = cols;
      cols += p;
    }
    int sidx = (int)sectors.size();
    sectors.push_back({kv.first, offset, rows, cols});
    sector_index[kv.first] = sidx;
    for (size_t j = 0; j < dtrees.size(); ++j) {
      for (size_t i = 0; i < ctrees.size(); ++i) {
        BlockInfo b;
        b.key.cod = ctrees[i];
        b.key.dom = dtrees[j];
        b.shape = rsh[i];
        b.shape.insert(b.shape.end(), csh[j].begin(), csh[j].end());
        int64_t st = 1;
        for (auto d : rsh[i]) {
          b.strides.push_back(st);
          st *= d;
        }
        st = rows;
        for (auto d : csh[j]) {
          b.strides.push_back(st);
          st *= d;
        }
        b.offset = offset + rstart[i] + rows * cstart[j];
        b.sector = sidx;
        index.emplace(b.key, (int)blocks.size());
        blocks.push_back(std::move(b));
      }
    }
    offset += rows * cols;
  }
  len = offset;
}

Leg HomSpace::leg_as_codomain(size_t axis) const {
  if (axis < cod.size()) return cod[axis];
  return dom[axis - cod.size()].dualized(rule);
}

std::shared_ptr<HomSpace> make_space(const Rule& rule, std::vector<Leg> cod, std::vector<Leg> dom) {
  auto s = std::make_shared<HomSpace>();
  s->rule = rule;
  s->cod = std::move(cod);
  s->dom = std::move(dom);
  if (s->rank() > TNT_MAX_RANK) throw Error("rank exceeds TNT_MAX_RANK");
  s->build();
  return s;
}

std::shared_ptr<HomSpace> permuted_space(const HomSpace& s, const std::vector<size_t>& p1,
                                         const std::vector<size_t>& p2) {
  std::vector<Leg> c, d;
  for (auto a : p1) c.push_back(s.leg_as_codomain(a));
  for (auto a : p2) d.push_back(s.leg_as_codomain(a).dualized(s.rule));
  return make_space(s.rule, c, d);
}

std::shared_ptr<HomSpace> adjoint_space(const HomSpace& s) { return make_space(s.rule, s.dom, s.cod); }

// ---------------------------------------------------------------------------
// Tree-pair operations
// ---------------------------------------------------------------------------
static Sid ext(const Tree& t, size_t k) {
  size_t n = t.unc.size();
  if (k == 0) return t.unc[0];
  if (k == n - 1) return t.coupled;
  return t.inner[k - 1];
}

using TreeTerms = std::vector<std::pair<Tree, double>>;

static void artin(const Rule& rule, const Tree& tree, size_t index, bool inverse, TreeTerms& out) {
  size_t n = tree.unc.size();
  Sid left = tree.unc[index], right = tree.unc[index + 1];
  Tree o = tree;
  std::swap(o.unc[index], o.unc[index + 1]);
  std::swap(o.isdual[index], o.isdual[index + 1]);
  Sid vac = rule.vacuum();
  if (left == vac || right == vac) {
    if (index > 0) o.inner[index - 1] = (left == vac) ? ext(o, index + 1) : ext(o, index - 1);
    out.emplace_back(std::move(o), 1.0);
    return;
  }
  if (index == 0) {
    Sid coupled = n > 2 ? tree.inner[0] : tree.coupled;
    double c = inverse ? rule.R(right, left, coupled) : rule.R(left, right, coupled);
    out.emplace_back(std::move(o), c);
    return;
  }
  Sid a = ext(tree, index - 1), b = left, c = ext(tree, index), d = right, e = ext(tree, index + 1);
  std::vector<Sid> ch;
  rule.channels(a, d, ch);
  for (Sid cp : ch) {
    if (!rule.nsym(cp, b, e)) continue;
    Tree t = o;
    t.inner[index - 1] = cp;
    double f = rule.F(d, a, b, e, cp, c);
    double coeff = inverse ? rule.R(d, c, e) * f * rule.R(d, a, cp)
                           : rule.R(c, d, e) * f * rule.R(a, d, cp);
    out.emplace_back(std::move(t), coeff);
  }
}

template <class K, class H>
static void accumulate(std::vector<std::pair<K, double>>& terms) {
  std::unordered_map<K, size_t, H> pos;
  std::vector<std::pair<K, double>> out;
  for (auto& kv : terms) {
    auto it = pos.find(kv.first);
    if (it == pos.end()) {
      pos.emplace(kv.first, out.size());
      out.push_back(std::move(kv));
    } else {
      out[it->second].second += kv.second;
    }
  }
  terms.swap(out);
}

struct TreeHash {
  size_t operator()(const Tree& t) const { return tree_hash(t); }
};

static std::vector<std::pair<size_t, bool>> braid_steps(const std::vector<size_t>& perm,
                                                        std::vector<size_t> lev) {
  size_t rank = perm.size();
  std::vector<size_t> work = perm;
  std::vector<std::pair<size_t, bool>> steps;
  for (size_t target = 0; target + 1 < rank; ++target) {
    size_t source = work[target];
    for (size_t index = source; index-- > target;) {
      steps.emplace_back(index, lev[index] > lev[index + 1]);
      std::swap(lev[index], lev[index + 1]);
    }
    for (size_t i = target + 1; i < rank; ++i)
      if (work[i] < source) ++work[i];
    work[target] = target;
  }
  return steps;
}

static PairKey bendright(const Rule& rule, const PairKey& p, double& coeff) {
  const Tree& f1 = p.cod;
  const Tree& f2 = p.dom;
  size_t n1 = f1.unc.size(), n2 = f2.unc.size();
  if (n1 == 0) throw Error("bendright needs a codomain leg");
  Sid b = f1.unc.back();
  bool bdual = f1.isdual.back();
  Sid c = f1.coupled;
  Sid a = n1 == 1 ? rule.vacuum() : (n1 == 2 ? f1.unc[0] : f1.inner.back());
  PairKey q;
  q.cod.unc.assign(f1.unc.begin(), f1.unc.end() - 1);
  q.cod.isdual.assign(f1.isdual.begin(), f1.isdual.end() - 1);
  q.cod.coupled = a;
  if (n1 > 2) q.cod.inner.assign(f1.inner.begin(), f1.inner.end() - 1);
  q.dom.unc = f2.unc;
  q.dom.unc.push_back(rule.dual(b));
  q.dom.isdual = f2.isdual;
  q.dom.isdual.push_back(!bdual);
  q.dom.coupled = a;
  q.dom.inner = f2.inner;
  if (n2 > 1) q.dom.inner.push_back(c);
  double k = rule.sqrtdim(c) / rule.sqrtdim(a) * rule.bsym(a, b, c);
  if (bdual) k *= rule.fs(rule.dual(b));
  coeff *= k;
  return q;
}

static PairKey bendleft(const Rule& rule, const PairKey& p, double& coeff) {
  PairKey sw{p.dom, p.cod};
  PairKey r = bendright(rule, sw, coeff);
  return PairKey{r.dom, r.cod};
}

static PairKey repartition(const Rule& rule, PairKey p, size_t target, double& coeff) {
  while (p.cod.unc.size() < target) p = bendleft(rule, p, coeff);
  while (p.cod.unc.size() > target) p = bendright(rule, p, coeff);
  return p;
}

Terms permute_pair(const Rule& rule, const PairKey& pair, const std::vector<size_t>& p1,
                   const std::vector<size_t>& p2) {
  size_t n1 = pair.cod.unc.size(), n2 = pair.dom.unc.size(), n = n1 + n2;
  bool ident = p1.size() == n1;
  for (size_t i = 0; ident && i < p1.size(); ++i) ident = p1[i] == i;
  for (size_t i = 0; ident && i < p2.size(); ++i) ident = p2[i] == n1 + i;
  if (ident) return {{pair, 1.0}};
  std::vector<size_t> perm(n);
  for (size_t pos = 0; pos < n; ++pos) {
    size_t axis = pos < p1.size() ? p1[pos] : p2[p2.size() - 1 - (pos - p1.size())];
    perm[pos] = axis < n1 ? axis : n2 + 2 * n1 - 1 - axis;
  }
  bool lin_ident = true;
  for (size_t i = 0; i < n; ++i) lin_ident = lin_ident && perm[i] == i;
  if (lin_ident) {
    double c = 1.0;
    PairKey q = repartition(rule, pair, p1.size(), c);
    return {{q, c}};
  }
  double c0 = 1.0;
  PairKey all = repartition(rule, pair, n, c0);
  std::vector<size_t> lev(n);
  for (size_t i = 0; i < n1; ++i) lev[i] = i;
  for (size_t i = 0; i < n2; ++i) lev[n1 + i] = n - 1 - i;
  auto steps = braid_steps(perm, lev);
  TreeTerms cur{{all.cod, 1.0}}, nxt;
  for (auto& st : steps) {
    nxt.clear();
    for (auto& tc : cur) {
      size_t before = nxt.size();
      artin(rule, tc.first, st.first, st.second, nxt);
      for (size_t i = before; i < nxt.size(); ++i) nxt[i].second *= tc.second;
    }
    accumulate<Tree, TreeHash>(nxt);
    cur.swap(nxt);
  }
  Terms out;
  for (auto& tc : cur) {
    double c = c0 * tc.second;
    PairKey q = repartition(rule, PairKey{tc.first, all.dom}, p1.size(), c);
    out.emplace_back(std::move(q), c);
  }
  accumulate<PairKey, PairKeyHash>(out);
  return out;
}

// ---------------------------------------------------------------------------
// Transform compile
// ---------------------------------------------------------------------------
void CompiledTransform::finalize() {
  desc.dst_len = dst->len;
  desc.src_len = src->len;
  desc.n_layouts = layouts.size();
  desc.layouts = layouts.data();
  desc.n_blocks = blocks.size();
  desc.blocks = blocks.data();
  desc.n_coeffs = coeffs.size();
  desc.coeffs = coeffs.data();
  desc.n_inactive = inactive.size();
  desc.inactive = inactive.data();
}

static tnt_layout make_layout(int64_t offset, const std::vector<int64_t>& shape,
                              const std::vector<int64_t>& strides) {
  tnt_layout l{};
  l.offset = offset;
  l.rank = (int32_t)shape.size();
  for (size_t i = 0; i < shape.size(); ++i) {
    l.shape[i] = shape[i];
    l.strides[i] = strides[i];
  }
  return l;
}

std::unique_ptr<CompiledTransform> compile_transform(const std::shared_ptr<HomSpace>& src,
                                                     bool adjoint_first,
                                                     const std::vector<size_t>& p1,
                                                     const std::vector<size_t>& p2) {
  const Rule& rule = src->rule;
  size_t n = src->rank();
  size_t nout = src->cod.size();
  if (p1.size() + p2.size() != n) throw Error("permutation rank mismatch");
  {
    std::vector<int> seen(n, 0);
    for (auto a : p1) { if (a >= n || seen[a]++) throw Error("invalid permutation"); }
    for (auto a : p2) { if (a >= n || seen[a]++) throw Error("invalid permutation"); }
  }
  // Logical source: src itself or its adjoint (codomain/domain swapped).
  std::shared_ptr<HomSpace> logical = adjoint_first ? adjoint_space(*src) : src;
  // logical axis i <-> stored axis adj[i]
  std::vector<size_t> adj(n);
  for (size_t i = 0; i < n; ++i) adj[i] = adjoint_first ? (i < n - nout ? nout + i : i - (n - nout)) : i;
  auto out = std::make_unique<CompiledTransform>();
  out->src = src;
  out->dst = permuted_space(*logical, p1, p2);
  const HomSpace& dst = *out->dst;
  std::vector<size_t> perm(p1);
  perm.insert(perm.end(), p2.begin(), p2.end());

  // group source blocks by uncoupled sectors (FusionTreeGroupKey), first-seen order
  struct GKey {
    std::vector<Sid> cu, du;
    std::vector<uint8_t> cd, dd;
    bool operator<(const GKey& o) const {
      return std::tie(cu, du, cd, dd) < std::tie(o.cu, o.du, o.cd, o.dd);
    }
  };
  std::map<GKey, size_t> gpos;
  std::vector<std::vector<int>> groups;
  auto logical_key = [&](int bi) {
    const PairKey& k = src->blocks[bi].key;
    return adjoint_first ? PairKey{k.dom, k.cod} : k;
  };
  for (int bi = 0; bi < (int)src->blocks.size(); ++bi) {
    PairKey k = logical_key(bi);
    GKey g{k.cod.unc, k.dom.unc, k.cod.isdual, k.dom.isdual};
    auto it = gpos.find(g);
    if (it == gpos.end()) {
      gpos.emplace(g, groups.size());
      groups.push_back({bi});
    } else {
      groups[it->second].push_back(bi);
    }
  }
  std::vector<uint8_t> touched(dst.blocks.size(), 0);
  auto src_layout = [&](int bi) {
    const BlockInfo& b = src->blocks[bi];
    std::vector<int64_t> shape(n), strides(n);
    for (size_t k = 0; k < n; ++k) {
      shape[k] = b.shape[adj[perm[k]]];
      strides[k] = b.strides[adj[perm[k]]];
    }
    return make_layout(b.offset, shape, strides);
  };
  auto dst_layout = [&](int di) {
    const BlockInfo& b = dst.blocks[di];
    return make_layout(b.offset, b.shape, b.strides);
  };
  std::vector<Terms> rows;
  for (auto& g : groups) {
    rows.clear();
    for (int bi : g) rows.push_back(permute_pair(rule, logical_key(bi), p1, p2));
    std::vector<PairKey> dkeys;
    std::unordered_map<PairKey, size_t, PairKeyHash> dpos;
    std::unordered_map<PairKey, int, PairKeyHash> seen;
    bool direct = true;
    for (auto& r : rows) {
      if (r.size() != 1 || seen.count(r[0].first)) direct = false;
      if (r.size() == 1) seen.emplace(r[0].first, 1);
      for (auto& kc : r) {
        if (!dpos.count(kc.first)) {
          dpos.emplace(kc.first, dkeys.size());
          dkeys.push_back(kc.first);
        }
      }
    }
    std::vector<int> didx;
    for (auto& k : dkeys) {
      auto it = dst.index.find(k);
      if (it == dst.index.end()) throw Error("destination fusion-tree pair missing from destination space");
      didx.push_back(it->second);
    }
    if (direct) {
      for (size_t i = 0; i < g.size(); ++i) {
        int di = dst.index.at(rows[i][0].first);
        touched[di] = 1;
        tnt_transform_block blk{};
        blk.kind = TNT_BLOCK_SINGLE;
        blk.dst_layout = (uint32_t)out->layouts.size();
        out->layouts.push_back(dst_layout(di));
        blk.dst_count = 1;
        blk.src_layout = (uint32_t)out->layouts.size();
        out->layouts.push_back(src_layout(g[i]));
        blk.src_count = 1;
        blk.coeff = (uint32_t)out->coeffs.size();
        out->coeffs.push_back(rows[i][0].second);
        out->blocks.push_back(blk);
        out->n_single++;
      }
    } else {
      tnt_transform_block blk{};
      blk.kind = TNT_BLOCK_MULTI;
      blk.dst_layout = (uint32_t)out->layouts.size();
      blk.dst_count = (uint32_t)didx.size();
      for (int di : didx) {
        touched[di] = 1;
        out->layouts.push_back(dst_layout(di));
      }
      blk.src_layout = (uint32_t)out->layouts.size();
      blk.src_count = (uint32_t)g.size();
      for (int bi : g) out->layouts.push_back(src_layout(bi));
      blk.coeff = (uint32_t)out->coeffs.size();
      size_t base = out->coeffs.size();
      out->coeffs.resize(base + didx.size() * g.size(), 0.0);
      for (size_t col = 0; col < rows.size(); ++col)
        for (auto& kc : rows[col]) out->coeffs[base + dpos.at(kc.first) * g.size() + col] += kc.second;
      int64_t elems = 1;
      for (auto d : src->blocks[g[0]].shape) elems *= d;
      out->recoupling_flops += 2.0 * double(elems) * double(g.size()) * double(didx.size());
      out->blocks.push_back(blk);
      out->n_multi++;
    }
  }
  for (size_t di = 0; di < dst.blocks.size(); ++di) {
    if (!touched[di]) {
      out->inactive.push_back((uint32_t)out->layouts.size());
      out->layouts.push_back(dst_layout((int)di));
    }
  }
  out->finalize();
  return out;
}

// ---------------------------------------------------------------------------
// Contraction compile (DynamicTree route)
// ---------------------------------------------------------------------------
std::unique_ptr<CompiledContract> compile_contract(const std::shared_ptr<HomSpace>& lhs,
                                                   const std::shared_ptr<HomSpace>& rhs,
                                                   const std::vector<size_t>& axes_l,
                                                   const std::vector<size_t>& axes_r,
                                                   const std::vector<size_t>& out_perm,
                                                   size_t out_nout, bool lhs_adj, bool rhs_adj) {
  if (lhs->rule.kind != rhs->rule.kind) throw Error("operands use different fusion rules");
  if (axes_l.size() != axes_r.size()) throw Error("contracted axis lists differ in length");
  size_t nl = lhs->rank(), nr = rhs->rank();
  size_t l_nout = lhs_adj ? lhs->dom.size() : lhs->cod.size();
  size_t r_nout = rhs_adj ? rhs->dom.size() : rhs->cod.size();
  std::vector<size_t> open_l, open_r;
  for (size_t i = 0; i < nl; ++i)
    if (std::find(axes_l.begin(), axes_l.end(), i) == axes_l.end()) open_l.push_back(i);
  for (size_t i = 0; i < nr; ++i)
    if (std::find(axes_r.begin(), axes_r.end(), i) == axes_r.end()) open_r.push_back(i);
  if (open_l.size() + axes_l.size() != nl || open_r.size() + axes_r.size() != nr)
    throw Error("contracted axes out of range or repeated");
  auto c = std::make_unique<CompiledContract>();
  auto is_ident = [](const std::vector<size_t>& a, const std::vector<size_t>& b, size_t nout) {
    if (a.size() != nout) return false;
    for (size_t i = 0; i < a.size(); ++i)
      if (a[i] != i) return false;
    for (size_t i = 0; i < b.size(); ++i)
      if (b[i] != nout + i) return false;
    return true;
  };
  // operand transforms
  std::shared_ptr<HomSpace> lt, rt;
  int lhs_op = TNT_OP_IDENTITY, rhs_op = TNT_OP_IDENTITY;
  if (is_ident(open_l, axes_l, l_nout)) {
    lt = lhs_adj ? adjoint_space(*lhs) : lhs;
    lhs_op = lhs_adj ? TNT_OP_ADJOINT : TNT_OP_IDENTITY;
  } else {
    c->lhs_t = compile_transform(lhs, lhs_adj, open_l, axes_l);
    lt = c->lhs_t->dst;
  }
  if (is_ident(axes_r, open_r, r_nout)) {
    rt = rhs_adj ? adjoint_space(*rhs) : rhs;
    rhs_op = rhs_adj ? TNT_OP_ADJOINT : TNT_OP_IDENTITY;
  } else {
    c->rhs_t = compile_transform(rhs, rhs_adj, axes_r, open_r);
    rt = c->rhs_t->dst;
  }
  if (!(lt->dom == rt->cod)) throw Error("contracted legs do not match (space mismatch)");
  c->core = make_space(lhs->rule, lt->cod, rt->dom);
  // stored-operand sector tables for the GEMM (adjoint operands address the stored matrix)
  const HomSpace& lstore = (lhs_op == TNT_OP_ADJOINT) ? *lhs : *lt;
  const HomSpace& rstore = (rhs_op == TNT_OP_ADJOINT) ? *rhs : *rt;
  for (auto& s : c->core->sectors) {
    auto li = lstore.sector_index.find(s.coupled);
    auto ri = rstore.sector_index.find(s.coupled);
    if (li == lstore.sector_index.end() || ri == rstore.sector_index.end()) {
      c->uncovered.push_back({s.offset, s.rows * s.cols});
      continue;
    }
    const SectorInfo& L = lstore.sectors[li->second];
    const SectorInfo& R = rstore.sectors[ri->second];
    int64_t lr = lhs_op == TNT_OP_ADJOINT ? L.cols : L.rows;
    int64_t lk = lhs_op == TNT_OP_ADJOINT ? L.rows : L.cols;
    int64_t rk = rhs_op == TNT_OP_ADJOINT ? R.cols : R.rows;
    int64_t rc = rhs_op == TNT_OP_ADJOINT ? R.rows : R.cols;
    if (lr != s.rows || rc != s.cols || lk != rk) throw Error("core sector dimensions disagree");
    if (lk == 0) {
      c->uncovered.push_back({s.offset, s.rows * s.cols});
      continue;
    }
    tnt_gemm_job j{(size_t)s.offset, (size_t)L.offset, (size_t)R.offset, (size_t)s.rows,
                   (size_t)lk, (size_t)s.cols};
    c->jobs.push_back(j);
    c->flops += 2.0 * double(s.rows) * double(lk) * double(s.cols);
  }
  // output transform
  size_t nopen = open_l.size() + open_r.size();
  if (out_perm.size() != nopen || out_nout > nopen) throw Error("output permutation rank mismatch");
  std::vector<size_t> q1(out_perm.begin(), out_perm.begin() + out_nout),
      q2(out_perm.begin() + out_nout, out_perm.end());
  if (is_ident(q1, q2, open_l.size())) {
    c->dst = c->core;
  } else {
    c->out_t = compile_transform(c->core, false, q1, q2);
    c->dst = c->out_t->dst;
  }
  auto& d = c->desc;
  d.lhs = c->lhs_t ? &c->lhs_t->desc : nullptr;
  d.rhs = c->rhs_t ? &c->rhs_t->desc : nullptr;
  d.out = c->out_t ? &c->out_t->desc : nullptr;
  d.n_jobs = c->jobs.size();
  d.jobs = c->jobs.data();
  d.lhs_len = lstore.len;
  d.rhs_len = rstore.len;
  d.core_len = c->core->len;
  d.n_uncovered = c->uncovered.size();
  d.uncovered = c->uncovered.data();
  d.lhs_op = lhs_op;
  d.rhs_op = rhs_op;
  return c;
}

}  // namespace tnth
