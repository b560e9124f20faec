// extern "C" surface of the host structure compiler (include/tenet_host.h).
#include <algorithm>
#include <cstring>
#include <string>

#include "../../../include/tenet_host.h"
#include "../common/batch_desc.hpp"
#include "../common/narrow_fusion.hpp"
#include "structure.hpp"

using namespace tnth;

struct tnth_space {
  std::shared_ptr<HomSpace> s;
};
struct tnth_transform {
  std::unique_ptr<CompiledTransform> t;
  tnth_space dst;
};
struct tnth_contract {
  std::unique_ptr<CompiledContract> c;
  tnth_space dst;
};

static thread_local std::string g_err;

template <class F>
static int guard(F&& f) {
  try {
    f();
    return TNT_OK;
  } catch (const Error& e) {
    g_err = e.msg;
    return TNT_E_INVALID;
  } catch (const std::exception& e) {
    g_err = e.what();
    return TNT_E_INVALID;
  }
}

static Rule rule_of(int r) {
  if (r < 0 || r > 3) throw Error("unknown rule");
  Rule rule;
  rule.kind = (RuleKind)r;
  return rule;
}

extern "C" {

const char* tnth_last_error(void) { return g_err.c_str(); }

uint64_t tnth_u1_id(int64_t q) { return u1_id(q); }
int64_t tnth_u1_charge(uint64_t id) { return u1_charge(id); }
uint64_t tnth_product_id(uint64_t l, uint64_t r) { return cantor(l, r); }
int tnth_product_split(uint64_t id, uint64_t* l, uint64_t* r) {
  Sid a, b;
  uncantor(id, a, b);
  *l = a;
  *r = b;
  return TNT_OK;
}

double tnth_symbol_f(int r, uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e, uint64_t f) {
  return rule_of(r).F(a, b, c, d, e, f);
}
double tnth_symbol_r(int r, uint64_t a, uint64_t b, uint64_t c) { return rule_of(r).R(a, b, c); }
double tnth_symbol_dim(int r, uint64_t a) { return rule_of(r).dim(a); }
uint64_t tnth_symbol_dual(int r, uint64_t a) { return rule_of(r).dual(a); }

int tnth_space_create(int rule, size_t n_cod, size_t n_dom, const size_t* n_sectors,
                      const uint64_t* ids, const int64_t* degs, const uint8_t* is_dual,
                      tnth_space** out) {
  return guard([&] {
    Rule r = rule_of(rule);
    std::vector<Leg> legs;
    size_t pos = 0;
    for (size_t l = 0; l < n_cod + n_dom; ++l) {
      std::vector<std::pair<Sid, int64_t>> p;
      for (size_t i = 0; i < n_sectors[l]; ++i, ++pos) {
        if (degs[pos] <= 0) throw Error("degeneracy must be positive");
        p.emplace_back(ids[pos], degs[pos]);
      }
      std::sort(p.begin(), p.end());
      Leg leg;
      for (size_t i = 0; i < p.size(); ++i) {
        if (i && p[i].first == p[i - 1].first) throw Error("sector declared more than once on a leg");
        leg.sectors.push_back(p[i].first);
        leg.degs.push_back(p[i].second);
      }
      leg.is_dual = is_dual ? is_dual[l] != 0 : false;
      legs.push_back(std::move(leg));
    }
    std::vector<Leg> c(legs.begin(), legs.begin() + n_cod), d(legs.begin() + n_cod, legs.end());
    *out = new tnth_space{make_space(r, c, d)};
  });
}

void tnth_space_destroy(tnth_space* s) { delete s; }
int tnth_space_rule(const tnth_space* s) { return (int)s->s->rule.kind; }
size_t tnth_space_nout(const tnth_space* s) { return s->s->cod.size(); }
size_t tnth_space_nin(const tnth_space* s) { return s->s->dom.size(); }
int64_t tnth_space_len(const tnth_space* s) { return s->s->len; }
size_t tnth_space_n_blocks(const tnth_space* s) { return s->s->blocks.size(); }

int tnth_space_block(const tnth_space* s, size_t i, int64_t* offset, int32_t* rank, int64_t* shape,
                     int64_t* strides, uint64_t* coupled) {
  if (i >= s->s->blocks.size()) {
    g_err = "block index out of range";
    return TNT_E_OOB;
  }
  const BlockInfo& b = s->s->blocks[i];
  if (offset) *offset = b.offset;
  if (rank) *rank = (int32_t)b.shape.size();
  for (size_t k = 0; k < b.shape.size(); ++k) {
    if (shape) shape[k] = b.shape[k];
    if (strides) strides[k] = b.strides[k];
  }
  if (coupled) *coupled = b.key.cod.coupled;
  return TNT_OK;
}

int tnth_space_block_trees(const tnth_space* s, size_t i, uint64_t* unc, uint64_t* ci, uint64_t* di) {
  if (i >= s->s->blocks.size()) {
    g_err = "block index out of range";
    return TNT_E_OOB;
  }
  const BlockInfo& b = s->s->blocks[i];
  size_t k = 0;
  for (auto x : b.key.cod.unc) unc[k++] = x;
  for (auto x : b.key.dom.unc) unc[k++] = x;
  if (ci)
    for (size_t j = 0; j < b.key.cod.inner.size(); ++j) ci[j] = b.key.cod.inner[j];
  if (di)
    for (size_t j = 0; j < b.key.dom.inner.size(); ++j) di[j] = b.key.dom.inner[j];
  return TNT_OK;
}

size_t tnth_space_n_sectors(const tnth_space* s) { return s->s->sectors.size(); }

int tnth_space_sector(const tnth_space* s, size_t i, uint64_t* coupled, int64_t* offset,
                      int64_t* rows, int64_t* cols, double* dim) {
  if (i >= s->s->sectors.size()) {
    g_err = "sector index out of range";
    return TNT_E_OOB;
  }
  const SectorInfo& x = s->s->sectors[i];
  if (coupled) *coupled = x.coupled;
  if (offset) *offset = x.offset;
  if (rows) *rows = x.rows;
  if (cols) *cols = x.cols;
  if (dim) *dim = s->s->rule.dim(x.coupled);
  return TNT_OK;
}

int tnth_space_leg(const tnth_space* s, size_t axis, size_t* n, uint64_t* ids, int64_t* degs,
                   uint8_t* is_dual) {
  const HomSpace& h = *s->s;
  if (axis >= h.rank()) {
    g_err = "axis out of range";
    return TNT_E_OOB;
  }
  const Leg& l = axis < h.cod.size() ? h.cod[axis] : h.dom[axis - h.cod.size()];
  if (n) *n = l.sectors.size();
  for (size_t i = 0; i < l.sectors.size(); ++i) {
    if (ids) ids[i] = l.sectors[i];
    if (degs) degs[i] = l.degs[i];
  }
  if (is_dual) *is_dual = l.is_dual;
  return TNT_OK;
}

int tnth_space_permuted(const tnth_space* s, const size_t* p1, size_t n1, const size_t* p2,
                        size_t n2, tnth_space** out) {
  return guard([&] {
    std::vector<size_t> a(p1, p1 + n1), b(p2, p2 + n2);
    *out = new tnth_space{permuted_space(*s->s, a, b)};
  });
}

int tnth_space_adjoint(const tnth_space* s, tnth_space** out) {
  return guard([&] { *out = new tnth_space{adjoint_space(*s->s)}; });
}

int tnth_space_equal(const tnth_space* a, const tnth_space* b) { return a->s->same(*b->s) ? 1 : 0; }

int tnth_transform_compile(const tnth_space* src, const size_t* p1, size_t n1, const size_t* p2,
                           size_t n2, tnth_transform** out) {
  return guard([&] {
    std::vector<size_t> a(p1, p1 + n1), b(p2, p2 + n2);
    auto t = compile_transform(src->s, false, a, b);
    auto* h = new tnth_transform;
    h->dst.s = t->dst;
    h->t = std::move(t);
    *out = h;
  });
}

void tnth_transform_destroy(tnth_transform* t) { delete t; }
const tnt_transform_desc* tnth_transform_desc(const tnth_transform* t) { return &t->t->desc; }
const tnth_space* tnth_transform_dst_space(const tnth_transform* t) { return &t->dst; }
int tnth_transform_counts(const tnth_transform* t, size_t* ns, size_t* nm) {
  *ns = t->t->n_single;
  *nm = t->t->n_multi;
  return TNT_OK;
}

int tnth_contract_compile(const tnth_space* lhs, const tnth_space* rhs, const size_t* axes_l,
                          const size_t* axes_r, size_t n_axes, const size_t* out_perm,
                          size_t out_nout, int lhs_adjoint, int rhs_adjoint, tnth_contract** out) {
  return guard([&] {
    std::vector<size_t> al(axes_l, axes_l + n_axes), ar(axes_r, axes_r + n_axes);
    size_t nopen = lhs->s->rank() + rhs->s->rank() - 2 * n_axes;
    std::vector<size_t> op(out_perm, out_perm + nopen);
    auto c = compile_contract(lhs->s, rhs->s, al, ar, op, out_nout, lhs_adjoint != 0,
                              rhs_adjoint != 0);
    auto* h = new tnth_contract;
    h->dst.s = c->dst;
    h->c = std::move(c);
    *out = h;
  });
}

void tnth_contract_destroy(tnth_contract* c) { delete c; }
const tnt_contract_desc* tnth_contract_desc(const tnth_contract* c) { return &c->c->desc; }
const tnth_space* tnth_contract_dst_space(const tnth_contract* c) { return &c->dst; }
double tnth_contract_flops(const tnth_contract* c) { return c->c->flops; }

struct tnth_batch {
  tnt_batch::BatchedContract b;
};

int tnth_contract_batch(const tnt_contract_desc* d, size_t copies, int shared, tnth_batch** out) {
  if (!d || !out) {
    g_err = "null argument";
    return TNT_E_INVALID;
  }
  auto* h = new tnth_batch;
  if (!tnt_batch::batch_contract(*d, copies, shared, h->b, g_err)) {
    delete h;
    *out = nullptr;
    return TNT_E_INVALID;
  }
  *out = h;
  return TNT_OK;
}

const tnt_contract_desc* tnth_batch_desc(const tnth_batch* b) { return &b->b.desc; }
void tnth_batch_destroy(tnth_batch* b) { delete b; }

struct tnth_fusion {
  tnt_fuse::NarrowFusion f;
};

int tnth_narrow_fusion(const tnt_contract_desc* d, tnth_fusion** out) {
  if (!d || !out) {
    g_err = "null argument";
    return TNT_E_INVALID;
  }
  auto* h = new tnth_fusion;
  if (!tnt_fuse::plan_narrow_fusion(d, h->f)) {
    g_err = h->f.reason;
    delete h;
    *out = nullptr;
    return TNT_E_UNSUPPORTED;
  }
  *out = h;
  return TNT_OK;
}
void tnth_fusion_destroy(tnth_fusion* f) { delete f; }
const tnt_transform_desc* tnth_fusion_desc(const tnth_fusion* f) { return &f->f.desc; }
size_t tnth_fusion_gather(const tnth_fusion* f, const int64_t** ptr, const int64_t** idx, const double** scale) {
  *ptr = f->f.gather_ptr.data();
  *idx = f->f.gather_idx.data();
  *scale = f->f.gather_scale.data();
  return f->f.gather_ptr.empty() ? 0 : f->f.gather_ptr.size() - 1;
}

struct tnth_bytes {
  std::vector<uint8_t> v;
};
const uint8_t* tnth_bytes_data(const tnth_bytes* b) { return b ? b->v.data() : nullptr; }
size_t tnth_bytes_size(const tnth_bytes* b) { return b ? b->v.size() : 0; }
void tnth_bytes_free(tnth_bytes* b) { delete b; }

int tnth_snapshot_tensor(const tnth_space* s, const double* payload, int64_t len, int adjoint, tnth_bytes** out) {
  return guard([&] {
    if (!s || !out || (!payload && len)) throw Error("null argument");
    *out = new tnth_bytes{encode_tensor_snapshot(*s->s, payload, len, adjoint != 0)};
  });
}

int tnth_snapshot_tensor_space(const uint8_t* bytes, size_t n, tnth_space** out, int* adjoint) {
  return guard([&] {
    if (!bytes || !out) throw Error("null argument");
    bool adj = false;
    auto sp = decode_tensor_space(bytes, n, adj);
    if (adjoint) *adjoint = adj ? 1 : 0;
    *out = new tnth_space{sp};
  });
}

int tnth_snapshot_tensor_payload(const uint8_t* bytes, size_t n, const tnth_space* s, double* payload, int64_t len) {
  return guard([&] {
    if (!bytes || !s || (!payload && len)) throw Error("null argument");
    decode_tensor_payload(bytes, n, *s->s, payload, len);
  });
}

int tnth_snapshot_leg(int rule, size_t n, const uint64_t* ids, const int64_t* degs, int is_dual, tnth_bytes** out) {
  return guard([&] {
    Rule r = rule_of(rule);
    if (!out || (n && (!ids || !degs))) throw Error("null argument");
    std::vector<std::pair<Sid, int64_t>> p;
    for (size_t i = 0; i < n; ++i) p.emplace_back(ids[i], degs[i]);
    std::sort(p.begin(), p.end());
    Leg leg;
    for (auto& x : p) {
      leg.sectors.push_back(x.first);
      leg.degs.push_back(x.second);
    }
    leg.is_dual = is_dual != 0;
    *out = new tnth_bytes{encode_leg_snapshot(r.kind, leg)};
  });
}

int tnth_snapshot_leg_decode(const uint8_t* bytes, size_t n, int* rule, size_t* n_sectors, uint64_t* ids,
                             int64_t* degs, int* is_dual) {
  return guard([&] {
    if (!bytes) throw Error("null argument");
    RuleKind k;
    Leg leg = decode_leg_snapshot(bytes, n, k);
    if (rule) *rule = (int)k;
    if (n_sectors) *n_sectors = leg.sectors.size();
    if (is_dual) *is_dual = leg.is_dual ? 1 : 0;
    for (size_t i = 0; i < leg.sectors.size(); ++i) {
      if (ids) ids[i] = leg.sectors[i];
      if (degs) degs[i] = leg.degs[i];
    }
  });
}

}  // extern "C"
