/*
 * tenet_host.h -- host-side structure compiler for the block-sparse symmetric
 * contraction path (no GPU needed).
 *
 * In TeNeT these steps live in the host compile layer and stay on the host:
 * the backend only consumes their outputs (compiled offsets, strides,
 * recoupling coefficients and GEMM job lists, as tnt_transform_desc /
 * tnt_contract_desc of tenet_hip.h).  This library restates that layer in
 * C++ (no Rust toolchain is assumed) to drive the backend end to end:
 *
 *   tnth_space_*         FusionTreeHomSpace + coupled-sector block layout
 *                        (tenet-core/src/fusion_space.rs:2211-2261, :3953-4104;
 *                        fusion_tree.rs:491-575, :1076-1131)
 *   tnth_transform_*     tree-pair permute coefficients (fusion_tree.rs:3278-3420,
 *                        :8088-8255, :8749-9040), group planning
 *                        (tenet-tensors/src/tree_transform/plan.rs:1489-1540,
 *                        :2107-2200) and structure compile
 *                        (tenet-operations/src/transform_structure.rs:587-776)
 *   tnth_contract_*      DynamicTree contraction artifact
 *                        (tenet-tensors/src/contract/dynamic.rs:807-817,
 *                        contract/fusion/plan.rs:264-278) with the canonical core
 *                        plan (tenet-operations/src/fusion_replay.rs:457-555)
 *   tnth_symbol_*        provider symbols (tenet-sectors/src/abelian.rs:723-780,
 *                        su2.rs:297-370, algebra.rs:552-625)
 */
#ifndef TENET_HOST_H
#define TENET_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "tenet_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

enum tnth_rule { TNTH_U1 = 0, TNTH_SU2 = 1, TNTH_Z2 = 2, TNTH_U1xZ2 = 3 };

typedef struct tnth_space tnth_space;
typedef struct tnth_transform tnth_transform;
typedef struct tnth_contract tnth_contract;

const char* tnth_last_error(void);

/* Sector ids (encodings of the reference providers). */
uint64_t tnth_u1_id(int64_t charge);
int64_t tnth_u1_charge(uint64_t id);
uint64_t tnth_product_id(uint64_t left, uint64_t right);
int tnth_product_split(uint64_t id, uint64_t* left, uint64_t* right);

/* Provider symbols (doubles; SectorIds as arguments). */
double tnth_symbol_f(int rule, uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e, uint64_t f);
double tnth_symbol_r(int rule, uint64_t a, uint64_t b, uint64_t c);
double tnth_symbol_dim(int rule, uint64_t a);
uint64_t tnth_symbol_dual(int rule, uint64_t a);

/* A HomSpace codomain <- domain.  Legs are given as concatenated (id, deg)
 * lists: leg l has n_sectors[l] entries starting at the running offset;
 * legs 0..n_cod-1 form the codomain, the rest the domain. */
int tnth_space_create(int rule, size_t n_cod, size_t n_dom, const size_t* n_sectors,
                      const uint64_t* sector_ids, const int64_t* degs, const uint8_t* is_dual,
                      tnth_space** out);
void tnth_space_destroy(tnth_space* space);
int tnth_space_rule(const tnth_space* space);
size_t tnth_space_nout(const tnth_space* space);
size_t tnth_space_nin(const tnth_space* space);
int64_t tnth_space_len(const tnth_space* space);
size_t tnth_space_n_blocks(const tnth_space* space);
/* Block i: offset, rank, shape and strides (arrays of TNT_MAX_RANK), coupled id. */
int tnth_space_block(const tnth_space* space, size_t i, int64_t* offset, int32_t* rank,
                     int64_t* shape, int64_t* strides, uint64_t* coupled);
/* Block i's tree labels: uncoupled (rank), codomain inner (max(nout-2,0)), domain
 * inner (max(nin-2,0)). */
int tnth_space_block_trees(const tnth_space* space, size_t i, uint64_t* uncoupled,
                           uint64_t* cod_inner, uint64_t* dom_inner);
size_t tnth_space_n_sectors(const tnth_space* space);
int tnth_space_sector(const tnth_space* space, size_t i, uint64_t* coupled, int64_t* offset,
                      int64_t* rows, int64_t* cols, double* dim);
/* Legs in the caller's format (see tnth_space_create); sizes first when arrays are NULL. */
int tnth_space_leg(const tnth_space* space, size_t axis, size_t* n, uint64_t* ids, int64_t* degs,
                   uint8_t* is_dual);
/* Destination space of permute(p1; p2). */
int tnth_space_permuted(const tnth_space* space, const size_t* p1, size_t n1, const size_t* p2,
                        size_t n2, tnth_space** out);
int tnth_space_adjoint(const tnth_space* space, tnth_space** out);
int tnth_space_equal(const tnth_space* a, const tnth_space* b);

/* Compiled permute src -> permuted(src; p1; p2). */
int tnth_transform_compile(const tnth_space* src, const size_t* p1, size_t n1, const size_t* p2,
                           size_t n2, tnth_transform** out);
void tnth_transform_destroy(tnth_transform* t);
const tnt_transform_desc* tnth_transform_desc(const tnth_transform* t);
const tnth_space* tnth_transform_dst_space(const tnth_transform* t);
/* Number of Single / Multi blocks (diagnostics). */
int tnth_transform_counts(const tnth_transform* t, size_t* n_single, size_t* n_multi);

/* Contraction: contract lhs axes_l[i] with rhs axes_r[i]; output axes
 * (open lhs..., open rhs...) permuted by out_perm, first out_nout in the
 * codomain.  lhs_adjoint/rhs_adjoint contract the adjoint of an operand
 * (lazy adjoint, MatrixOp::Adjoint). */
int tnth_contract_compile(const tnth_space* lhs, const tnth_space* rhs, const size_t* axes_l,
                          const size_t* axes_r, size_t n_axes, const size_t* out_perm,
                          size_t out_nout, int lhs_adjoint, int rhs_adjoint, tnth_contract** out);
void tnth_contract_destroy(tnth_contract* c);
const tnt_contract_desc* tnth_contract_desc(const tnth_contract* c);
const tnth_space* tnth_contract_dst_space(const tnth_contract* c);
/* Algorithmic FP64 flops of the core GEMMs (sum 2 m k n) and of the recoupling GEMMs. */
double tnth_contract_flops(const tnth_contract* c);

/* Fused route of a narrow contraction (every core job k <= 32, n <= 16; no
 * reference counterpart -- the reference runs such contractions as transform,
 * per-sector GEMM, transform).  The backend plans the same route inside
 * tnt_contract_upload; this entry point exposes the plan for inspection and
 * CPU checks: a transform from the lhs payload to the output payload made of
 * Multi groups whose coefficients are sums of rhs terms,
 * U[t] = sum over ptr[t] <= p < ptr[t+1] of scale[p] * rhs[idx[p]]
 * (tnth_fusion_gather returns the number of coefficients; ptr has one more
 * entry).  Returns TNT_E_UNSUPPORTED, with the failed condition in
 * tnth_last_error(), when the contraction is not eligible. */
/* Stacked (batched) contraction: `copies` independent instances of the
 * contraction d in one descriptor, for one tnt_contract call that runs them
 * all with the launches of one (the serving loop's cross-request batching;
 * no reference counterpart -- the reference contracts one tensor pair per
 * call, tenet-tensors/src/contract/dynamic.rs:1241-1530).  shared: 0 = every
 * copy reads the same lhs payload (given once, its transform run once), the
 * rhs and the destination are stacked (copy c at c * length); 1 = the rhs is
 * shared; 2 = nothing is shared.  The result owns its tables. */
typedef struct tnth_batch tnth_batch;
int tnth_contract_batch(const tnt_contract_desc* d, size_t copies, int shared, tnth_batch** out);
const tnt_contract_desc* tnth_batch_desc(const tnth_batch* b);
void tnth_batch_destroy(tnth_batch* b);

typedef struct tnth_fusion tnth_fusion;
int tnth_narrow_fusion(const tnt_contract_desc* d, tnth_fusion** out);
void tnth_fusion_destroy(tnth_fusion* f);
const tnt_transform_desc* tnth_fusion_desc(const tnth_fusion* f);
size_t tnth_fusion_gather(const tnth_fusion* f, const int64_t** ptr, const int64_t** idx, const double** scale);

/* ---- snapshots: TeNeT's v1 semantic wire format ----------------------------
 * Replace TensorMap::to_bytes_with / from_bytes_with and GradedSpace's pair
 * (tenet/src/typed/serialization.rs:912-957, :1569-1640, :1700-1740) for dense
 * f64 tensors, with a fixed codec per rule: provider keys "u1", "su2", "z2",
 * "u1xz2"; labels = charge as LE i32 (U1), one byte 2j (SU2), one byte parity
 * (Z2), charge then parity (U1xZ2) -- the reference test codecs
 * (tenet/tests/typed_serialization.rs:63-131) for U1 and SU2.
 * Encoders return a byte buffer owned by the caller (tnth_bytes_free). */
typedef struct tnth_bytes tnth_bytes;
const uint8_t* tnth_bytes_data(const tnth_bytes* b);
size_t tnth_bytes_size(const tnth_bytes* b);
void tnth_bytes_free(tnth_bytes* b);
/* Tensor snapshot of a host copy of the payload (adjoint != 0: a lazy adjoint). */
int tnth_snapshot_tensor(const tnth_space* space, const double* payload, int64_t len, int adjoint,
                         tnth_bytes** out);
/* Decode in two steps: the space (and adjoint flag), then the payload into a
 * caller buffer of tnth_space_len(space) doubles (untouched on failure). */
int tnth_snapshot_tensor_space(const uint8_t* bytes, size_t n, tnth_space** out, int* adjoint);
int tnth_snapshot_tensor_payload(const uint8_t* bytes, size_t n, const tnth_space* space,
                                 double* payload, int64_t len);
/* One graded space (leg); ids as in tnth_space_create. */
int tnth_snapshot_leg(int rule, size_t n, const uint64_t* ids, const int64_t* degs, int is_dual,
                      tnth_bytes** out);
/* Sizes first (ids, degs NULL), then the sectors sorted by id. */
int tnth_snapshot_leg_decode(const uint8_t* bytes, size_t n, int* rule, size_t* n_sectors,
                             uint64_t* ids, int64_t* degs, int* is_dual);

#ifdef __cplusplus
}
#endif

#endif /* TENET_HOST_H */
