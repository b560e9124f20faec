// Versioned semantic snapshots of spaces and tensors in TeNeT's v1 wire
// format (tenet/src/typed/serialization.rs:17-26 constants, :301-356 framing,
// :673-708 legs, :710-731 spaces, :733-775 trees and block keys, :813-870 dense
// bodies, :912-957 tensors), so that environments written by the device
// backend can be read by TeNeT and the other way round.
//
// The wire format leaves sector-label bytes and provider keys to a codec; this
// file fixes one codec per built-in rule, the same as the reference's own test
// codecs where one exists (tenet/tests/typed_serialization.rs:63-131):
//   U1     key "u1",    label = charge as little-endian i32
//   SU2    key "su2",   label = one byte 2j
//   Z2     key "z2",    label = one byte parity
//   U1xZ2  key "u1xz2", label = charge (LE i32) then one parity byte
//
// Layout (all integers little-endian, lengths u64):
//   "TENETTS\0" | u16 version = 1 | u8 kind (1 space, 2 tensor) |
//   u8 scalar (0 none, 1 f64) | blob provider key
//   space snapshot:  leg
//   tensor snapshot: u8 repr (1 dense | 3 adjoint followed by 1 dense) |
//                    u64 n_cod | u64 n_dom | legs | u64 n_blocks | blocks
//   leg:   u8 dual | u64 n | n x (blob label, u64 degeneracy), sorted by label bytes
//   block: blob key | u64 rank | rank x u64 dim | u64 count | count x f64,
//          values in logical order (first axis fastest), blocks sorted by key bytes
//   key:   codomain tree | domain tree;  tree: blob coupled | u64 n + labels of
//          the uncoupled sectors | u64 n + labels of the inner lines |
//          u64 n + u64 vertex labels (all 1: multiplicity-free)
// Decoding checks framing and declared counts before allocating, rejects
// unknown versions, duplicate, unsorted or missing blocks and trailing bytes,
// and never returns a partially filled payload as success.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../../include/tenet_host.h"
#include "structure.hpp"

namespace tnth {

namespace {

constexpr char kMagic[8] = {'T', 'E', 'N', 'E', 'T', 'T', 'S', '\0'};
constexpr uint16_t kVersion = 1;
constexpr uint8_t kKindSpace = 1, kKindTensor = 2;
constexpr uint8_t kScalarNone = 0, kScalarF64 = 1;
constexpr uint8_t kReprDense = 1, kReprAdjoint = 3;
// Resource ceilings on untrusted counts (checked before anything is allocated).
constexpr uint64_t kMaxRank = 64, kMaxSectorsPerLeg = 1u << 20, kMaxBlocks = 1u << 24;
constexpr uint64_t kMaxBlob = 1u << 20;

using Bytes = std::vector<uint8_t>;

const char* provider_key(RuleKind k) {
  switch (k) {
    case RuleKind::U1: return "u1";
    case RuleKind::SU2: return "su2";
    case RuleKind::Z2: return "z2";
    default: return "u1xz2";
  }
}

void put_u64(Bytes& o, uint64_t v) {
  for (int i = 0; i < 8; ++i) o.push_back(uint8_t(v >> (8 * i)));
}
void put_blob(Bytes& o, const Bytes& b) {
  put_u64(o, b.size());
  o.insert(o.end(), b.begin(), b.end());
}
void put_f64(Bytes& o, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  put_u64(o, u);
}

Bytes encode_label(RuleKind k, Sid id) {
  Bytes b;
  auto i32 = [&](int64_t q) {
    if (q < INT32_MIN || q > INT32_MAX) throw Error("U(1) charge outside the i32 label range");
    const uint32_t u = uint32_t(int32_t(q));
    for (int i = 0; i < 4; ++i) b.push_back(uint8_t(u >> (8 * i)));
  };
  switch (k) {
    case RuleKind::U1: i32(u1_charge(id)); break;
    case RuleKind::SU2:
      if (id > 255) throw Error("SU(2) label 2j outside one byte");
      b.push_back(uint8_t(id));
      break;
    case RuleKind::Z2: b.push_back(uint8_t(id)); break;
    default: {
      Sid l, r;
      uncantor(id, l, r);
      i32(u1_charge(l));
      b.push_back(uint8_t(r));
    }
  }
  return b;
}

Sid decode_label(RuleKind k, const uint8_t* p, uint64_t n) {
  auto i32 = [&](const uint8_t* q) {
    uint32_t u = 0;
    for (int i = 0; i < 4; ++i) u |= uint32_t(q[i]) << (8 * i);
    return int64_t(int32_t(u));
  };
  switch (k) {
    case RuleKind::U1:
      if (n != 4) throw Error("invalid U(1) label");
      return u1_id(i32(p));
    case RuleKind::SU2:
      if (n != 1) throw Error("invalid SU(2) label");
      return p[0];
    case RuleKind::Z2:
      if (n != 1 || p[0] > 1) throw Error("invalid Z2 label");
      return p[0];
    default:
      if (n != 5 || p[4] > 1) throw Error("invalid U(1)xZ2 label");
      return cantor(u1_id(i32(p)), p[4]);
  }
}

void put_header(Bytes& o, uint8_t kind, uint8_t scalar, RuleKind k) {
  o.insert(o.end(), kMagic, kMagic + 8);
  o.push_back(uint8_t(kVersion & 0xff));
  o.push_back(uint8_t(kVersion >> 8));
  o.push_back(kind);
  o.push_back(scalar);
  const char* key = provider_key(k);
  put_blob(o, Bytes(key, key + std::strlen(key)));
}

void put_leg(Bytes& o, RuleKind k, const Leg& leg) {
  o.push_back(leg.is_dual ? 1 : 0);
  std::vector<std::pair<Bytes, int64_t>> e;
  for (size_t i = 0; i < leg.sectors.size(); ++i) e.emplace_back(encode_label(k, leg.sectors[i]), leg.degs[i]);
  std::sort(e.begin(), e.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  put_u64(o, e.size());
  for (auto& x : e) {
    put_blob(o, x.first);
    put_u64(o, uint64_t(x.second));
  }
}

void put_tree(Bytes& o, RuleKind k, const Tree& t) {
  put_blob(o, encode_label(k, t.coupled));
  put_u64(o, t.unc.size());
  for (Sid s : t.unc) put_blob(o, encode_label(k, s));
  put_u64(o, t.inner.size());
  for (Sid s : t.inner) put_blob(o, encode_label(k, s));
  const size_t nv = t.unc.empty() ? 0 : t.unc.size() - 1;
  put_u64(o, nv);
  for (size_t i = 0; i < nv; ++i) put_u64(o, 1);
}

struct Reader {
  const uint8_t* p;
  size_t n, pos = 0;
  const uint8_t* take(uint64_t k, const char* what) {
    if (k > n - pos) throw Error(std::string("truncated snapshot (") + what + ")");
    const uint8_t* r = p + pos;
    pos += k;
    return r;
  }
  uint8_t u8(const char* what) { return *take(1, what); }
  uint64_t u64(const char* what) {
    const uint8_t* q = take(8, what);
    uint64_t v = 0;
    for (int i = 0; i < 8; ++i) v |= uint64_t(q[i]) << (8 * i);
    return v;
  }
  std::pair<const uint8_t*, uint64_t> blob(const char* what) {
    const uint64_t k = u64(what);
    if (k > kMaxBlob) throw Error(std::string("blob longer than the limit (") + what + ")");
    return {take(k, what), k};
  }
};

RuleKind read_header(Reader& r, uint8_t kind, uint8_t scalar) {
  const uint8_t* m = r.take(8, "magic");
  if (std::memcmp(m, kMagic, 8) != 0) throw Error("not a TeNeT snapshot (bad magic)");
  const uint8_t* v = r.take(2, "version");
  const uint16_t ver = uint16_t(v[0] | (v[1] << 8));
  if (ver != kVersion) throw Error("unsupported snapshot version " + std::to_string(ver));
  if (r.u8("kind") != kind) throw Error("snapshot holds another kind of object");
  if (r.u8("scalar") != scalar) throw Error("snapshot scalar type is not the requested one");
  auto key = r.blob("provider key");
  const std::string s(reinterpret_cast<const char*>(key.first), key.second);
  for (int k = 0; k < 4; ++k)
    if (s == provider_key(RuleKind(k))) return RuleKind(k);
  throw Error("unknown provider key '" + s + "'");
}

Leg read_leg(Reader& r, RuleKind k) {
  Leg leg;
  const uint8_t d = r.u8("leg duality");
  if (d > 1) throw Error("invalid leg duality flag");
  leg.is_dual = d != 0;
  const uint64_t n = r.u64("leg sectors");
  if (n > kMaxSectorsPerLeg) throw Error("too many sectors on a leg");
  std::vector<std::pair<Sid, int64_t>> e;
  Bytes prev;
  for (uint64_t i = 0; i < n; ++i) {
    auto b = r.blob("sector label");
    Bytes cur(b.first, b.first + b.second);
    if (i && !(prev < cur)) throw Error("leg sectors not in canonical order");
    prev = cur;
    const uint64_t deg = r.u64("degeneracy");
    if (deg == 0 || deg > (uint64_t(1) << 40)) throw Error("invalid degeneracy");
    e.emplace_back(decode_label(k, b.first, b.second), int64_t(deg));
  }
  std::sort(e.begin(), e.end());
  for (auto& x : e) {
    leg.sectors.push_back(x.first);
    leg.degs.push_back(x.second);
  }
  return leg;
}

Tree read_tree(Reader& r, RuleKind k, const std::vector<Leg>& legs) {
  Tree t;
  auto c = r.blob("coupled");
  t.coupled = decode_label(k, c.first, c.second);
  const uint64_t nu = r.u64("uncoupled");
  if (nu != legs.size()) throw Error("tree has the wrong number of uncoupled sectors");
  for (uint64_t i = 0; i < nu; ++i) {
    auto b = r.blob("uncoupled");
    t.unc.push_back(decode_label(k, b.first, b.second));
    t.isdual.push_back(legs[i].is_dual ? 1 : 0);
  }
  const uint64_t ni = r.u64("inner lines");
  if (ni != (nu >= 2 ? nu - 2 : 0)) throw Error("tree has the wrong number of inner lines");
  for (uint64_t i = 0; i < ni; ++i) {
    auto b = r.blob("inner line");
    t.inner.push_back(decode_label(k, b.first, b.second));
  }
  const uint64_t nv = r.u64("vertices");
  if (nv != (nu >= 1 ? nu - 1 : 0)) throw Error("tree has the wrong number of vertices");
  for (uint64_t i = 0; i < nv; ++i)
    if (r.u64("vertex") != 1) throw Error("vertex multiplicity label other than 1");
  return t;
}

struct Parsed {
  RuleKind rule;
  bool adjoint = false;
  std::vector<Leg> cod, dom;
  size_t blocks_pos = 0;  // reader position of the block count
};

Parsed read_tensor_head(Reader& r) {
  Parsed p;
  p.rule = read_header(r, kKindTensor, kScalarF64);
  uint8_t repr = r.u8("representation");
  if (repr == kReprAdjoint) {
    p.adjoint = true;
    repr = r.u8("representation");
  }
  if (repr != kReprDense) throw Error("only dense f64 tensor snapshots are supported");
  const uint64_t nc = r.u64("codomain rank"), nd = r.u64("domain rank");
  if (nc + nd > kMaxRank || nc > kMaxRank || nd > kMaxRank) throw Error("rank exceeds the limit");
  for (uint64_t i = 0; i < nc; ++i) p.cod.push_back(read_leg(r, p.rule));
  for (uint64_t i = 0; i < nd; ++i) p.dom.push_back(read_leg(r, p.rule));
  p.blocks_pos = r.pos;
  return p;
}

}  // namespace

Bytes encode_leg_snapshot(RuleKind k, const Leg& leg) {
  Bytes o;
  put_header(o, kKindSpace, kScalarNone, k);
  put_leg(o, k, leg);
  return o;
}

Leg decode_leg_snapshot(const uint8_t* bytes, size_t n, RuleKind& k) {
  Reader r{bytes, n};
  k = read_header(r, kKindSpace, kScalarNone);
  Leg leg = read_leg(r, k);
  if (r.pos != n) throw Error("trailing bytes after the snapshot");
  return leg;
}

Bytes encode_tensor_snapshot(const HomSpace& s, const double* data, int64_t len, bool adjoint) {
  if (len != s.len) throw Error("payload length does not match the space");
  const RuleKind k = s.rule.kind;
  Bytes o;
  put_header(o, kKindTensor, kScalarF64, k);
  if (adjoint) o.push_back(kReprAdjoint);
  o.push_back(kReprDense);
  put_u64(o, s.cod.size());
  put_u64(o, s.dom.size());
  for (const Leg& l : s.cod) put_leg(o, k, l);
  for (const Leg& l : s.dom) put_leg(o, k, l);
  std::vector<std::pair<Bytes, size_t>> keys;
  for (size_t b = 0; b < s.blocks.size(); ++b) {
    Bytes key;
    put_tree(key, k, s.blocks[b].key.cod);
    put_tree(key, k, s.blocks[b].key.dom);
    keys.emplace_back(std::move(key), b);
  }
  std::sort(keys.begin(), keys.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  put_u64(o, keys.size());
  for (auto& kb : keys) {
    const BlockInfo& blk = s.blocks[kb.second];
    put_blob(o, kb.first);
    put_u64(o, blk.shape.size());
    uint64_t count = 1;
    for (int64_t d : blk.shape) {
      put_u64(o, uint64_t(d));
      count *= uint64_t(d);
    }
    put_u64(o, count);
    std::vector<int64_t> idx(blk.shape.size(), 0);
    for (uint64_t e = 0; e < count; ++e) {
      int64_t off = blk.offset;
      for (size_t a = 0; a < idx.size(); ++a) off += idx[a] * blk.strides[a];
      put_f64(o, data[off]);
      for (size_t a = 0; a < idx.size(); ++a) {  // first axis fastest
        if (++idx[a] < blk.shape[a]) break;
        idx[a] = 0;
      }
    }
  }
  return o;
}

std::shared_ptr<HomSpace> decode_tensor_space(const uint8_t* bytes, size_t n, bool& adjoint) {
  Reader r{bytes, n};
  Parsed p = read_tensor_head(r);
  adjoint = p.adjoint;
  Rule rule;
  rule.kind = p.rule;
  return make_space(rule, p.cod, p.dom);
}

void decode_tensor_payload(const uint8_t* bytes, size_t n, const HomSpace& s, double* out, int64_t len) {
  Reader r{bytes, n};
  Parsed p = read_tensor_head(r);
  if (p.rule != s.rule.kind || p.cod != s.cod || p.dom != s.dom) throw Error("snapshot space differs from the given space");
  if (len != s.len) throw Error("payload length does not match the space");
  const uint64_t nb = r.u64("block count");
  if (nb > kMaxBlocks) throw Error("block count exceeds the limit");
  if (nb != s.blocks.size()) throw Error("snapshot holds " + std::to_string(nb) + " blocks, the space has " +
                                         std::to_string(s.blocks.size()));
  std::vector<uint8_t> seen(s.blocks.size(), 0);
  Bytes prev;
  // values are staged so that a failure leaves `out` untouched
  std::vector<double> stage(size_t(len), 0.0);
  for (uint64_t b = 0; b < nb; ++b) {
    auto kb = r.blob("block key");
    Bytes key(kb.first, kb.first + kb.second);
    if (b && !(prev < key)) throw Error("blocks duplicated or not in canonical order");
    prev = key;
    Reader kr{kb.first, size_t(kb.second)};
    PairKey pk;
    pk.cod = read_tree(kr, p.rule, s.cod);
    pk.dom = read_tree(kr, p.rule, s.dom);
    if (kr.pos != kr.n) throw Error("trailing bytes in a block key");
    auto it = s.index.find(pk);
    if (it == s.index.end()) throw Error("block key is not a block of the space");
    const BlockInfo& blk = s.blocks[it->second];
    if (seen[it->second]) throw Error("duplicate block");
    seen[it->second] = 1;
    const uint64_t rank = r.u64("block rank");
    if (rank != blk.shape.size()) throw Error("block rank differs from the space");
    uint64_t count = 1;
    for (uint64_t a = 0; a < rank; ++a) {
      const uint64_t d = r.u64("block dim");
      if (d != uint64_t(blk.shape[a])) throw Error("block shape differs from the space");
      count *= d;
    }
    if (r.u64("value count") != count) throw Error("block value count differs from its shape");
    const uint8_t* v = r.take(8 * count, "block values");
    std::vector<int64_t> idx(rank, 0);
    for (uint64_t e = 0; e < count; ++e) {
      int64_t off = blk.offset;
      for (size_t a = 0; a < rank; ++a) off += idx[a] * blk.strides[a];
      uint64_t u = 0;
      for (int i = 0; i < 8; ++i) u |= uint64_t(v[8 * e + i]) << (8 * i);
      std::memcpy(&stage[size_t(off)], &u, 8);
      for (size_t a = 0; a < rank; ++a) {
        if (++idx[a] < blk.shape[a]) break;
        idx[a] = 0;
      }
    }
  }
  if (r.pos != n) throw Error("trailing bytes after the snapshot");
  std::copy(stage.begin(), stage.end(), out);
}

}  // namespace tnth
