// Host-side structure compiler: sectors, fusion trees, coupled-sector layout,
// tree-pair recoupling and compilation into the flat descriptors of
// tenet_hip.h.  Cold path (runs once per structure); no GPU code here.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/tenet_hip.h"

namespace tnth {

using Sid = uint64_t;

enum class RuleKind : int { U1 = 0, SU2 = 1, Z2 = 2, U1xZ2 = 3 };

// Provider symbols of the reference's built-in multiplicity-free rules.
struct Rule {
  RuleKind kind = RuleKind::U1;
  bool unique() const { return kind != RuleKind::SU2; }
  Sid vacuum() const;
  Sid dual(Sid a) const;
  void channels(Sid a, Sid b, std::vector<Sid>& out) const;  // ascending ids
  int nsym(Sid a, Sid b, Sid c) const;
  double F(Sid a, Sid b, Sid c, Sid d, Sid e, Sid f) const;
  double R(Sid a, Sid b, Sid c) const;
  double dim(Sid a) const;
  double sqrtdim(Sid a) const;
  double fs(Sid a) const;
  double bsym(Sid a, Sid b, Sid c) const;  // algebra.rs:614-624
};

Sid u1_id(int64_t q);
int64_t u1_charge(Sid id);
Sid cantor(Sid l, Sid r);
void uncantor(Sid id, Sid& l, Sid& r);
double sixj_doubled(int64_t j1, int64_t j2, int64_t j3, int64_t j4, int64_t j5, int64_t j6);

struct Leg {
  std::vector<Sid> sectors;  // sorted by id
  std::vector<int64_t> degs;
  bool is_dual = false;
  int64_t deg(Sid s) const;
  Leg dualized(const Rule& rule) const;
  bool operator==(const Leg& o) const {
    return sectors == o.sectors && degs == o.degs && is_dual == o.is_dual;
  }
};

struct Tree {
  std::vector<Sid> unc;
  std::vector<uint8_t> isdual;
  Sid coupled = 0;
  std::vector<Sid> inner;
  bool operator==(const Tree& o) const {
    return coupled == o.coupled && unc == o.unc && inner == o.inner && isdual == o.isdual;
  }
};

struct PairKey {
  Tree cod, dom;
  bool operator==(const PairKey& o) const { return cod == o.cod && dom == o.dom; }
};

struct PairKeyHash {
  size_t operator()(const PairKey& k) const;
};

struct BlockInfo {
  PairKey key;
  int64_t offset = 0;
  std::vector<int64_t> shape, strides;
  int sector = 0;
};

struct SectorInfo {
  Sid coupled;
  int64_t offset, rows, cols;
};

struct HomSpace {
  Rule rule;
  std::vector<Leg> cod, dom;
  std::vector<BlockInfo> blocks;
  std::unordered_map<PairKey, int, PairKeyHash> index;
  std::vector<SectorInfo> sectors;
  std::unordered_map<Sid, int> sector_index;
  int64_t len = 0;

  size_t rank() const { return cod.size() + dom.size(); }
  void build();  // enumerate trees and lay out blocks
  Leg leg_as_codomain(size_t axis) const;
  bool same(const HomSpace& o) const { return rule.kind == o.rule.kind && cod == o.cod && dom == o.dom; }
};

std::shared_ptr<HomSpace> make_space(const Rule& rule, std::vector<Leg> cod, std::vector<Leg> dom);
std::shared_ptr<HomSpace> permuted_space(const HomSpace& s, const std::vector<size_t>& p1,
                                         const std::vector<size_t>& p2);
std::shared_ptr<HomSpace> adjoint_space(const HomSpace& s);

// Coefficient expansion of permute(pair; p1, p2).
using Terms = std::vector<std::pair<PairKey, double>>;
Terms permute_pair(const Rule& rule, const PairKey& pair, const std::vector<size_t>& p1,
                   const std::vector<size_t>& p2);

// Owning storage behind a tnt_transform_desc.
struct CompiledTransform {
  std::shared_ptr<HomSpace> src, dst;
  std::vector<tnt_layout> layouts;
  std::vector<tnt_transform_block> blocks;
  std::vector<double> coeffs;
  std::vector<uint32_t> inactive;
  tnt_transform_desc desc{};
  size_t n_single = 0, n_multi = 0;
  double recoupling_flops = 0.0;  // 2 * element_count * src_count * dst_count per Multi
  void finalize();
};

// Compile src -> permuted(src; p1; p2).  With adjoint_first the source is
// the (real) adjoint of `src` (lazy adjoint operand).
std::unique_ptr<CompiledTransform> compile_transform(const std::shared_ptr<HomSpace>& src,
                                                     bool adjoint_first,
                                                     const std::vector<size_t>& p1,
                                                     const std::vector<size_t>& p2);

struct CompiledContract {
  std::shared_ptr<HomSpace> dst, core;
  std::unique_ptr<CompiledTransform> lhs_t, rhs_t, out_t;
  std::vector<tnt_gemm_job> jobs;
  std::vector<tnt_region> uncovered;
  tnt_contract_desc desc{};
  double flops = 0.0;
};

std::unique_ptr<CompiledContract> compile_contract(const std::shared_ptr<HomSpace>& lhs,
                                                   const std::shared_ptr<HomSpace>& rhs,
                                                   const std::vector<size_t>& axes_l,
                                                   const std::vector<size_t>& axes_r,
                                                   const std::vector<size_t>& out_perm,
                                                   size_t out_nout, bool lhs_adj, bool rhs_adj);

struct Error : std::exception {
  std::string msg;
  explicit Error(std::string m) : msg(std::move(m)) {}
  const char* what() const noexcept override { return msg.c_str(); }
};

// Snapshots in TeNeT's v1 wire format (snapshot.cpp).
std::vector<uint8_t> encode_leg_snapshot(RuleKind k, const Leg& leg);
Leg decode_leg_snapshot(const uint8_t* bytes, size_t n, RuleKind& k);
std::vector<uint8_t> encode_tensor_snapshot(const HomSpace& s, const double* data, int64_t len, bool adjoint);
std::shared_ptr<HomSpace> decode_tensor_space(const uint8_t* bytes, size_t n, bool& adjoint);
void decode_tensor_payload(const uint8_t* bytes, size_t n, const HomSpace& s, double* out, int64_t len);


}  // namespace tnth
