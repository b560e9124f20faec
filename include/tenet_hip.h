/*
 * tenet_hip.h -- MI355X (gfx950) dense backend for TeNeT's block-sparse
 * symmetric contraction path: tree-transform (recoupling) replay and
 * per-coupled-sector block GEMM, behind a plain C ABI.
 *
 * This is the drop-in boundary.  Each entry point replaces one seam of the
 * reference (paths relative to the TeNeT repository root):
 *
 *   tnt_gemm_batch        DenseExecutor::matmul_batch_axpby_into /
 *                         matmul_batch_axpby_with_ops_into
 *                         (tenet-dense/src/executor.rs:217-312); also the
 *                         per-job StorageGemm::matmul_range_into seam
 *                         (tenet-operations/src/fusion_replay.rs:1302-1345,
 *                         tenet-operations/src/cuda.rs:53-113)
 *   tnt_plan_upload       compiled TreeTransformStructure handed to a
 *                         storage-aware backend
 *                         (tenet-operations/src/transform_structure.rs:82,
 *                         :1343-1360, :1753-1759)
 *   tnt_tree_transform    TreeTransformBackend::tree_transform_structure_into_raw
 *                         and ..._overwrite_into_raw
 *                         (tenet-operations/src/replay_backend.rs:28-176, :547-623)
 *   tnt_contract_upload / FusionContractPlan + FusionBlockContractPlan replay:
 *   tnt_contract          execute_dynamic_tree_execution_artifact_impl
 *                         (tenet-tensors/src/contract/dynamic.rs:1241-1530) and
 *                         FusionBlockContractPlan::execute_direct_on_storage_prezeroed
 *                         (tenet-operations/src/fusion_replay.rs:1195-1266)
 *   tnt_weighted_dot      weighted_inner_cuda (tenet/src/typed.rs:11210-11262)
 *   tnt_buf_*             CudaDenseStorage upload/download/zero
 *                         (tenet-dense/src/cuda_adapter.rs:80-218)
 *   tnt_ctx_*             CudaDenseContext (tenet-dense/src/cuda_adapter.rs:80-107);
 *                         one context = one leased executor (tenet/src/runtime.rs:802-840)
 *   tnt_graph_*           no reference counterpart (the reference launches one
 *                         GEMM per sector); batches a whole plan into a hipGraph.
 *
 * Conventions (kept from the reference):
 *   - all payloads are flat column-major element buffers; offsets are element
 *     offsets; leading dimension of a job operand = its row count
 *     (tenet-dense/src/executor.rs:8-13, :510-515);
 *   - destinations of one batch are pairwise disjoint, so jobs run in any order;
 *   - validate before mutate: every shape/offset check runs before any kernel;
 *   - device mismatch is an error, never a transfer; there is no host fallback.
 *
 * Every function returns 0 on success or a negative TNT_E* code; the message
 * of the last error of a context is returned by tnt_last_error().
 * A context is single-threaded (like a leased DenseExecutor); use one context
 * per host thread.
 */
#ifndef TENET_HIP_H
#define TENET_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TNT_MAX_RANK 8

enum tnt_status {
  TNT_OK = 0,
  TNT_E_INVALID = -1,
  TNT_E_UNSUPPORTED = -2,
  TNT_E_SHAPE = -3,
  TNT_E_OOB = -4,
  TNT_E_DEVICE_MISMATCH = -5,
  TNT_E_HIP = -6,
  TNT_E_OOM = -7
};

/* Payload element types: DenseRead / DenseWrite::{F64, F32, C32, C64}
 * (tenet-dense/src/executor.rs:228-285).  FP64 payloads are accepted by every
 * entry point (the reference's device subset is FP64,
 * tenet-dense/src/cuda_adapter.rs:108-140).  FP32 and complex payloads
 * (TNT_C32: two FP32, TNT_C64: two FP64 per element, real part first) are
 * accepted by tnt_buf_alloc / _wrap / _upload / _download / _zero / _info /
 * _free and by tnt_gemm_batch / tnt_gemm_batch_complex; lengths and offsets
 * count elements of the payload type.  Every other entry point returns
 * TNT_E_UNSUPPORTED for them. */
enum tnt_dtype { TNT_F64 = 0, TNT_F32 = 1, TNT_C32 = 2, TNT_C64 = 3 };

/* MatrixOp {Identity, Transpose, Adjoint} (tenet-dense/src/executor.rs:24-31),
 * same order.  On real payloads Transpose and Adjoint are the same operation;
 * on complex payloads Adjoint is the conjugate transpose. */
enum tnt_op { TNT_OP_IDENTITY = 0, TNT_OP_TRANSPOSE = 1, TNT_OP_ADJOINT = 2 };

/* Payload arithmetic of the block GEMM. */
enum tnt_precision {
  TNT_PREC_F64 = 0,        /* FP64 MFMA, FP64 accumulate */
  TNT_PREC_F32 = 1,        /* operands rounded to FP32, FP32 MFMA, FP32 accumulate */
  TNT_PREC_SPLIT_BF16 = 2, /* Ozaki split into BF16 slices, BF16 MFMA, exact FP32 chunk sums, FP64 accumulate */
  TNT_PREC_SPLIT_I8 = 3    /* Ozaki split into INT8 slices, INT8 MFMA, exact INT32 level sums, FP64 accumulate */
};

typedef struct tnt_ctx tnt_ctx;
typedef struct tnt_buf tnt_buf;
typedef struct tnt_plan tnt_plan;
typedef struct tnt_artifact tnt_artifact;
typedef struct tnt_graph tnt_graph;
typedef struct tnt_event tnt_event;

/* DenseGemmBatchJob (tenet-dense/src/executor.rs:15-22): column-major
 * dst[rows x cols] (+)= op(lhs)[rows x contracted] * op(rhs)[contracted x cols]. */
typedef struct {
  size_t dst_offset;
  size_t lhs_offset;
  size_t rhs_offset;
  size_t rows;
  size_t contracted;
  size_t cols;
} tnt_gemm_job;

/* One grouped-GEMM job over explicit device pointers (several contractions in
 * one launch).  Column-major with explicit leading dimensions; op (a tnt_op)
 * applies to the stored matrices.  rows, contracted, cols and the leading
 * dimensions must lie in [0, 2^31 - 1]. */
typedef struct {
  double* dst;
  const double* lhs;
  const double* rhs;
  int64_t rows, contracted, cols;
  int64_t ldd, ldl, ldr;
  int32_t lhs_op, rhs_op;
} tnt_gemm_ptr_job;

/* One strided layout of a TreeTransformLayout table
 * (tenet-operations/src/transform_structure.rs:1753-1759): `shape` is given in
 * DESTINATION axis order for both roles, `strides` are the element strides of
 * the block it addresses in its own payload. */
typedef struct {
  int64_t offset;
  int32_t rank;
  int32_t reserved;
  int64_t shape[TNT_MAX_RANK];
  int64_t strides[TNT_MAX_RANK];
} tnt_layout;

enum tnt_block_kind { TNT_BLOCK_SINGLE = 0, TNT_BLOCK_MULTI = 1 };

/* TreeTransformBlock (transform_structure.rs:1343-1360).
 * Single: dst layout `dst_layout`, src layout `src_layout`, coefficient
 *         coeffs[coeff].
 * Multi:  dst layouts dst_layout .. +dst_count, src layouts src_layout ..
 *         +src_count, row-major U[dst, src] at coeffs[coeff]; all layouts of
 *         one Multi block have the same element count. */
typedef struct {
  int32_t kind;
  uint32_t dst_layout;
  uint32_t dst_count;
  uint32_t src_layout;
  uint32_t src_count;
  uint32_t coeff;
} tnt_transform_block;

typedef struct {
  int64_t dst_len;                 /* required destination payload length */
  int64_t src_len;                 /* required source payload length */
  size_t n_layouts;
  const tnt_layout* layouts;
  size_t n_blocks;
  const tnt_transform_block* blocks;
  size_t n_coeffs;
  const double* coeffs;
  size_t n_inactive;               /* destination layouts written by no block */
  const uint32_t* inactive;        /* (scaled by beta / zeroed on overwrite) */
} tnt_transform_desc;

/* A contiguous region of a payload. */
typedef struct {
  int64_t offset;
  int64_t len;
} tnt_region;

/* A DynamicTree contraction artifact (contract/dynamic.rs:807-817):
 * dst = alpha * OUT( CORE( LHS(lhs), RHS(rhs) ) ) + beta * dst.
 * A NULL transform means that operand (or the core result) is used as is.
 * Core jobs address the transformed operands and the core destination. */
typedef struct {
  const tnt_transform_desc* lhs;
  const tnt_transform_desc* rhs;
  const tnt_transform_desc* out;
  size_t n_jobs;
  const tnt_gemm_job* jobs;
  int64_t lhs_len, rhs_len;        /* lengths of the (transformed) core operands */
  int64_t core_len;                /* length of the core destination */
  size_t n_uncovered;              /* core-destination regions written by no job */
  const tnt_region* uncovered;
  int32_t lhs_op, rhs_op;
} tnt_contract_desc;

/* ---- context -------------------------------------------------------------- */
int tnt_ctx_create(int device, tnt_ctx** out);
int tnt_ctx_destroy(tnt_ctx* ctx);
const char* tnt_last_error(const tnt_ctx* ctx);
int tnt_ctx_synchronize(tnt_ctx* ctx);
/* The context's HIP stream (hipStream_t), for interop and event timing. */
int tnt_ctx_stream(tnt_ctx* ctx, void** stream);
/* Device HBM bytes held by the context's pools. */
int tnt_ctx_pool_bytes(const tnt_ctx* ctx, size_t* bytes);

/* ---- device buffers (flat payloads) --------------------------------------- */
int tnt_buf_alloc(tnt_ctx* ctx, size_t len, int dtype, tnt_buf** out);
/* Wrap caller-owned device memory (not freed by tnt_buf_free). */
int tnt_buf_wrap(tnt_ctx* ctx, void* device_ptr, size_t len, int dtype, tnt_buf** out);
int tnt_buf_free(tnt_ctx* ctx, tnt_buf* buf);
int tnt_buf_upload(tnt_ctx* ctx, tnt_buf* buf, const void* host, size_t len);
int tnt_buf_download(tnt_ctx* ctx, const tnt_buf* buf, void* host, size_t len);
int tnt_buf_zero(tnt_ctx* ctx, tnt_buf* buf);
/* Seeded fill over the flat payload: normal == 0 gives the reference's
 * uniform draw in [-1, 1) (Fill::Rand, tenet/src/typed.rs:2086-2092, bitwise
 * identical); normal != 0 gives N(0,1) by Box-Muller on the same stream. */
int tnt_buf_fill_random(tnt_ctx* ctx, tnt_buf* buf, uint64_t seed, int normal);
int tnt_buf_info(const tnt_buf* buf, void** device_ptr, size_t* len, int* dtype, int* device);

/* ---- batched block GEMM ---------------------------------------------------- */
/* DenseExecutor::matmul_batch_axpby_with_ops_into (tenet-dense/src/executor.rs:292-312;
 * matmul_batch_axpby_into, :217-287, is lhs_op = rhs_op = Identity).
 * runs: optional strided-batch run partition (executor.rs:473-488); it is a
 * scheduling hint only and may be NULL.  lhs_op / rhs_op are tnt_op values
 * (MatrixOp); any other value is TNT_E_INVALID.  dst, lhs and rhs must share
 * one payload type (mixed types are TNT_E_INVALID):
 *   F64 / C64: every precision;
 *   F32 / C32: TNT_PREC_F64 (FP64 arithmetic, result rounded to FP32 once) or
 *              TNT_PREC_F32 (FP32 MFMA); split precisions TNT_E_UNSUPPORTED.
 * Complex jobs run as real m x 2k x 2n block GEMMs on FP64 planes of the
 * operands ([Re A | Im A] . [[Re B, Im B], [-Im B, Re B]]) in a grow-only
 * workspace of the context (2mk + 4kn + 2mn doubles over the batch), with
 * contracted and column counts up to 2^30 - 1.  Every check runs before
 * anything is queued.  tnt_gemm_batch passes real alpha and beta. */
int tnt_gemm_batch(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                   const tnt_gemm_job* jobs, size_t n_jobs, const size_t* runs, size_t n_runs,
                   int lhs_op, int rhs_op, double alpha, double beta, int precision);
/* The same with DenseScalar::{C32, C64} alpha and beta (tenet-dense/src/scalar.rs:5-11):
 * D = alpha op(A) op(B) + beta D in complex arithmetic, D not read when
 * beta = 0.  On real payloads alpha_im and beta_im must be 0 (else
 * TNT_E_INVALID). */
int tnt_gemm_batch_complex(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                           const tnt_gemm_job* jobs, size_t n_jobs, const size_t* runs, size_t n_runs,
                           int lhs_op, int rhs_op, double alpha_re, double alpha_im, double beta_re,
                           double beta_im, int precision);
/* Grouped GEMM over explicit pointers: all jobs in one persistent launch. */
int tnt_gemm_grouped(tnt_ctx* ctx, const tnt_gemm_ptr_job* jobs, size_t n_jobs, double alpha,
                     double beta, int precision);

/* ---- tree-transform (recoupling) replay ----------------------------------- */
int tnt_plan_upload(tnt_ctx* ctx, const tnt_transform_desc* desc, tnt_plan** out);
int tnt_plan_destroy(tnt_ctx* ctx, tnt_plan* plan);
/* dst = alpha * T(src) + beta * dst, or dst = alpha * T(src) when overwrite != 0
 * (inactive destination layouts are then zeroed, never read). */
int tnt_tree_transform(tnt_ctx* ctx, const tnt_plan* plan, tnt_buf* dst, const tnt_buf* src,
                       double alpha, double beta, int overwrite);

/* ---- contraction artifact -------------------------------------------------- */
int tnt_contract_upload(tnt_ctx* ctx, const tnt_contract_desc* desc, tnt_artifact** out);
int tnt_contract_destroy(tnt_ctx* ctx, tnt_artifact* art);
/* Rebuild the artifact's GEMM tile table with one tile configuration (see
 * tnt_set_tile_config for the codes; -1 = table default).  The autotuner
 * measures each configuration on the artifact and keeps the fastest. */
int tnt_contract_set_tile_config(tnt_ctx* ctx, tnt_artifact* art, int cfg);
/* Execution route chosen at upload (no reference counterpart).  A narrow
 * contraction -- every core job with k <= 32 contracted and n <= 16 output
 * columns, Single-only transforms -- runs as ONE recoupling pass from the lhs
 * payload to the destination whose coefficients are gathered from the rhs
 * payload, instead of transform + GEMM + transform (HBM-bound either way; the
 * fused pass moves each element once).  TENET_AMD_FUSE_NARROW=0 in the
 * environment at upload keeps the general route.
 * A tiny contraction -- at most 16 Mi flops, every payload at most 2^20
 * elements, Single-only transforms, no core element feeding two destination
 * elements -- runs its FP64 calls as ONE launch: the operand transforms are
 * folded into the GEMM's operand loads (element maps built at upload) and the
 * output transform into its stores (TNT_ROUTE_TINY_GATHERED); other
 * precisions keep the general route.  TENET_AMD_TINY_FLOPS=0 at upload keeps
 * the general route. */
enum tnt_route { TNT_ROUTE_TRANSFORM_GEMM = 0, TNT_ROUTE_FUSED_NARROW = 1, TNT_ROUTE_TINY_GATHERED = 2 };
int tnt_contract_route(const tnt_artifact* art, int* route);
/* Stream order: all work of the call is ordered on the context stream
 * (tnt_ctx_stream).  When both operand transforms move at least 64 MiB (or,
 * for TNT_PREC_SPLIT_I8 without an output transform, one of them does), one
 * transform runs on a second stream of the context, forked from and joined
 * back into the context stream by events inside the call, so work queued on
 * the context stream after the call still sees the whole result.  Under
 * graph capture, per-kernel timing and the shadow check it runs on the
 * context stream alone (TENET_AMD_FORK_TRANSFORMS=0: always). */
int tnt_contract(tnt_ctx* ctx, const tnt_artifact* art, tnt_buf* dst, const tnt_buf* lhs,
                 const tnt_buf* rhs, double alpha, double beta, int precision);

/* ---- reductions ------------------------------------------------------------ */
/* out = sum_r weights[r] * sum_{i in region r} a[i] * b[i]  (FP64, deterministic). */
int tnt_weighted_dot(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                     const double* weights, size_t n_regions, double* out);
/* Stream-ordered tnt_weighted_dot: the result is written to out[out_index] on
 * the device, no host synchronisation (the FP64 inner products of a device
 * Krylov loop, KrylovVector::dot_real, tenet-krylov/src/vector.rs:13). */
int tnt_weighted_dot_async(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                           const double* weights, size_t n_regions, tnt_buf* out, size_t out_index);
/* tnt_weighted_dot_async with no device allocation at all: no counterpart in
 * the reference.  For callers that cannot allocate (code under hipGraph
 * capture, tight serving and Krylov loops, harnesses that forbid hipMalloc).
 * The chunk tables and partial sums live in a fixed 1 MiB device array per
 * GPU (8 bytes per region + 28 per chunk of 65536 elements: up to about
 * 37000 chunks, 2.4 G elements); larger region lists return
 * TNT_E_UNSUPPORTED.  The tables are staged from host memory the context
 * holds, so nothing is allocated per call.  The result equals
 * tnt_weighted_dot_async's bit for bit.  Every call on one GPU shares the
 * array: calls on one context are ordered by its stream, calls that overlap
 * from two contexts of one GPU race. */
int tnt_weighted_dot_async_noalloc(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                                   const double* weights, size_t n_regions, tnt_buf* out, size_t out_index);
/* y = alpha * x + beta * y over whole payloads. */
int tnt_axpby(tnt_ctx* ctx, tnt_buf* y, const tnt_buf* x, double alpha, double beta);

/* ---- device scalars of a Krylov loop ---------------------------------------- */
/* The scalar recurrences of CG / Lanczos kept on the device, so that a loop
 * iteration queues its matvec, dots (tnt_weighted_dot_async into a scalar
 * buffer), scalar steps and vector updates without waiting, and reads its
 * scalars back once (the loop of tenet-krylov/src/cg.rs:64-190 over
 * KrylovVector::dot_real / axpy / scale, vector.rs:2-20, with the
 * synchronisation moved out of every dot).  No other reference counterpart. */
/* y = a x + b y, a = alpha_factor * s[alpha_index] (alpha_factor when the
 * index is negative), b likewise; beta_factor == 0 writes a x without reading
 * y, alpha_factor == 0 writes b y without reading x; nothing is written when
 * guard_index >= 0 and s[guard_index] == 0. */
int tnt_axpby_dev(tnt_ctx* ctx, tnt_buf* y, const tnt_buf* x, const tnt_buf* scalars, int64_t alpha_index,
                  double alpha_factor, int64_t beta_index, double beta_factor, int64_t guard_index);
enum tnt_scalar_code {
  TNT_SCALAR_SET = 0,      /* s[out] = imm */
  TNT_SCALAR_DIV = 1,      /* s[out] = s[a] / s[b] */
  TNT_SCALAR_COPY = 2,     /* s[out] = s[a] */
  TNT_SCALAR_RSQRT = 3,    /* s[out] = 1 / sqrt(s[a]) */
  TNT_SCALAR_NEG = 4,      /* s[out] = -s[a] */
  TNT_SCALAR_CG_GUARD = 5  /* s[out] = 1 if s[a] (curvature) and s[b] (step) are finite and, when
                              imm != 0, s[a] > 0; else 0 */
};
typedef struct {
  int32_t op, out, a, b;
  double imm;
} tnt_scalar_op;
/* Runs up to 8 scalar steps in order on one device thread (stream-ordered). */
int tnt_scalars(tnt_ctx* ctx, tnt_buf* scalars, const tnt_scalar_op* ops, size_t n_ops);

/* ---- completion events (serving loop) -------------------------------------- */
/* Record on the context stream; query returns 1 when all work queued before the
 * record has completed, 0 while it runs.  No reference counterpart. */
int tnt_event_create(tnt_ctx* ctx, tnt_event** out);
int tnt_event_record(tnt_ctx* ctx, tnt_event* ev);
int tnt_event_query(tnt_ctx* ctx, const tnt_event* ev);
int tnt_event_destroy(tnt_ctx* ctx, tnt_event* ev);

/* ---- kernel timing ---------------------------------------------------------- */
/* When on, every recoupling replay (kind 0), every block-GEMM table run (kind
 * 1 + precision: all launches of one call, split pre-passes included) and
 * every FP64/FP32 GEMM kernel launch (kind 8 + 8 * precision + tile shape, shapes
 * as in tnt_set_tile_config) is bracketed by HIP events on the context stream;
 * stats report launches, total milliseconds and total algorithmic work (bytes
 * for kind 0, flops otherwise).
 * No reference counterpart (the reference's profiles are host wall-clock,
 * tenet-operations/src/profile.rs:16-59). */
int tnt_ctx_set_timing(tnt_ctx* ctx, int on);
/* Debug shadow write counts (no reference counterpart; the reference relies
 * on its disjoint-destination invariant, tenet-dense/src/executor.rs:8-13).
 * When on, every recoupling replay, FP64/FP32 block-GEMM table run (with the
 * uncovered core regions) and tiny-route call is followed by a pass that
 * counts, per destination element, the writes of the work decomposition that
 * ran (plan items, GEMM tiles, regions, output-map stores): `writes` sums the
 * counts, `doubles` counts elements written more than once, `expected` sums
 * the writes the compiled plans prescribe, `passes` the checked
 * destinations, all since the last tnt_ctx_set_shadow(ctx, 1).  A correct
 * decomposition gives writes == expected and doubles == 0.  Not during graph
 * capture; allocates a 4-byte counter per destination element. */
int tnt_ctx_set_shadow(tnt_ctx* ctx, int on);
int tnt_ctx_shadow_stats(tnt_ctx* ctx, uint64_t* writes, uint64_t* expected, uint64_t* doubles, uint64_t* passes);
int tnt_ctx_kernel_stats(tnt_ctx* ctx, int kind, size_t* launches, double* total_ms, double* total_work);
int tnt_ctx_reset_stats(tnt_ctx* ctx);

/* ---- tile-shape autotuner --------------------------------------------------- */
/* Force one GEMM tile configuration for the tables this context builds after
 * the call (0: 128x128, 1: 64x64, 2: 128x64, 3: 64x128, 4: narrow-output
 * vector kernel, 5: 128x128 on 8 waves, 6: 32x32 per wave with operands in
 * registers (tiny jobs; also takes narrow jobs of <= 256 rows); -1: per-job
 * choice).  Narrow outputs otherwise keep the vector kernel.  Used by the
 * autotuner sweep; no reference counterpart.  Per context, like every knob
 * below: contexts leased to different threads do not share them. */
int tnt_set_tile_config(tnt_ctx* ctx, int cfg);
/* Install the tuned per-shape table: entries (m_max, n_max, cfg) scanned in
 * order, first match wins; n = 0 clears it. */
int tnt_set_tile_table(tnt_ctx* ctx, const int64_t* m_max, const int64_t* n_max, const int32_t* cfg, size_t n);

/* Number of digit planes per operand of TNT_PREC_SPLIT_BF16 and
 * TNT_PREC_SPLIT_I8 (2..6, default 5: relative error ~1e-12 (BF16) and
 * ~1e-11 (INT8) on random data; 4 trades accuracy for speed).  No reference
 * counterpart. */
int tnt_set_split_slices(tnt_ctx* ctx, int slices);

/* ---- test hooks --------------------------------------------------------------- */
/* Failure injection for tests of the error paths (no reference counterpart;
 * the reference's error handling is mirrored, runtime.rs:561-620,
 * fusion_replay.rs:610-631).  TNT_FAULT_ALLOC: the context's pool allocation
 * after `after` more successful ones fails with TNT_E_OOM (no HIP call is
 * made); TNT_FAULT_LAUNCH: the status check after the `after`-th next batch of
 * kernel launches reports TNT_E_HIP (the kernels themselves are queued
 * normally); TNT_FAULT_NONE clears both.  after = -1 disarms one kind. */
enum tnt_fault { TNT_FAULT_NONE = 0, TNT_FAULT_ALLOC = 1, TNT_FAULT_LAUNCH = 2 };
int tnt_ctx_inject_fault(tnt_ctx* ctx, int kind, int after);
/* Entries of the context's table cache (content-addressed device copies of
 * per-call job / region tables) kept before least recently used entries not
 * pinned by a live graph are evicted (default 512; tests use small caps). */
int tnt_ctx_set_table_cache_cap(tnt_ctx* ctx, size_t entries);

/* ---- hipGraph batching ----------------------------------------------------- */
/* Everything queued on the context between begin and end is recorded, not
 * run.  Every table a recorded call needs must exist before capture (run the
 * call once first); nothing is allocated or uploaded while capturing.  While
 * any graph of a context is alive, the memory it recorded is kept: the
 * context's cached tables used by the graph are pinned, and workspaces that
 * grow or tables that are rebuilt retire their old memory until the last
 * graph is destroyed.  Buffers, plans and artifacts the caller owns must
 * outlive the graphs that use them. */
int tnt_graph_begin(tnt_ctx* ctx);
int tnt_graph_end(tnt_ctx* ctx, tnt_graph** out);
int tnt_graph_launch(tnt_ctx* ctx, const tnt_graph* graph);
int tnt_graph_destroy(tnt_ctx* ctx, tnt_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* TENET_HIP_H */
