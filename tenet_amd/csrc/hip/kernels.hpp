// Internal device-side formats and kernel launchers of the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tnt_dev {

constexpr int kMaxRank = 8;

// Pointers in the global address space.  A pointer rebuilt from an integer
// address is generic (FLAT) to the compiler; FLAT loads count against lgkmcnt
// as well as vmcnt, so every wait for an LDS read would also wait for all
// global loads in flight (a GEMM's prefetch of its next K slab).  Kernels that
// take byte addresses convert them with gmem().
template <class T>
using gptr = T __attribute__((address_space(1)))*;
template <class T>
__device__ __forceinline__ gptr<T> gmem(uintptr_t addr) {
  return (gptr<T>)addr;
}
constexpr int kMaxMultiSrc = 32;  // source blocks per Multi group handled in registers
constexpr int kSingleUnroll = 8;  // elements per thread of a Single work item
constexpr int kSingleItem = 256 * kSingleUnroll;  // plane elements per Single work item
constexpr int64_t kMultiItemWork = 8192;  // element updates per Multi work item
constexpr int64_t kMinPlanItems = 1024;   // items per class a small plan is cut into

// XK_TRANSPOSE: a Single block whose source is contiguous along another axis
// than the destination (dims[0] destination-contiguous, dims[1]
// source-contiguous), recoupled through 64 x 64 LDS tiles.
enum XKind : int32_t { XK_SINGLE = 0, XK_MULTI = 1, XK_SCALE = 2, XK_TRANSPOSE = 3 };
constexpr int kTransposeTile = 64;

// One normalised transform block: dims in destination order (extent-1 axes
// dropped, axes sorted by destination stride, contiguous runs fused).
struct XBlock {
  int32_t kind;
  int32_t rank;     // >= 1; dims[0] is the destination-fastest axis
  int32_t ndst;     // destination layouts (Single: 1)
  int32_t nsrc;     // source layouts (Single: 1, Scale: 0)
  int32_t lay;      // first layout in the layout table: ndst destination layouts, then nsrc source
  int32_t coeff;    // index of the coefficient (Single) or of row-major U[dst, src] (Multi)
  int64_t dims[kMaxRank];
};

struct XLayout {
  int64_t offset;
  int64_t strides[kMaxRank];
};

// A work item: plane elements [e_begin, e_begin + e_count) (plane = dims 0
// and 1) for outer indices [o_begin, o_begin + o_count) over dims 2..rank-1.
struct XItem {
  uint32_t block;
  uint32_t o_count;
  uint64_t o_begin;
  uint32_t e_begin;
  uint32_t e_count;
};

// A Single item of one outer index, flattened at upload: base offsets of the
// item's plane and the plane strides, so that the kernel reads one record per
// item (no block / layout lookups, no outer-index division).
struct XFlat {
  int64_t od, os;          // destination / source offset of plane element (0, 0)
  int64_t d0, d1, s0, s1;  // plane strides (dims 0 and 1)
  double c;                // block coefficient (alpha applied in the kernel)
  uint32_t n0;             // plane extent along dims[0]
  uint32_t e_begin, e_count;
  uint32_t pad;
};

struct XformArgs {
  const XBlock* blocks;
  const XLayout* layouts;
  const XItem* items;
  const XFlat* flat;  // records of the XC_FLAT items, in item order
  const double* coeffs;
  double* dst;
  const double* src;
  double alpha;
  double beta;
  int overwrite;
  // coefficients gathered per item instead of read from coeffs (the fused
  // narrow route): coefficient c = sum over gptr[c] <= p < gptr[c + 1] of
  // gscale[p] * gsrc[gidx[p]] (launch_xform_multi_gather only)
  const int64_t* gptr = nullptr;
  const int64_t* gidx = nullptr;
  const double* gscale = nullptr;
  const double* gsrc = nullptr;
};

// Work-item classes of a plan: Single and Scale items, Multi items by source
// count (<= 4, <= 16, <= kMaxMultiSrc), transposed Single tiles, and Single
// items of one outer index (XC_FLAT, flattened records).  One kernel per
// class, except that XC_FLAT and XC_SINGLE share one launch.  A transpose
// item is one 64 x 64 tile: e_begin = tile index in the plane (x tiles
// fastest), o_begin = outer index.
enum XClass : int32_t {
  XC_SINGLE = 0,
  XC_MULTI4 = 1,
  XC_MULTI16 = 2,
  XC_MULTI32 = 3,
  XC_TRANSPOSE = 4,
  XC_FLAT = 5,
  XC_COUNT = 6
};
__host__ __device__ inline int xclass_of(int kind, int nsrc) {
  if (kind == XK_TRANSPOSE) return XC_TRANSPOSE;
  if (kind != XK_MULTI) return XC_SINGLE;
  return nsrc <= 4 ? XC_MULTI4 : nsrc <= 16 ? XC_MULTI16 : XC_MULTI32;
}
// Items [begin, begin + n) of one class (not XC_SINGLE / XC_FLAT).
void launch_xform(const XformArgs& a, int cls, uint32_t begin, uint32_t n, hipStream_t s);
// The Single class in one launch: n_flat flattened records (XformArgs::flat),
// then the XC_SINGLE items [begin, begin + n).
void launch_xform_single(const XformArgs& a, uint32_t begin, uint32_t n, uint32_t n_flat, hipStream_t s);
// All items of a plan in one launch (any class).
void launch_xform_any(const XformArgs& a, uint32_t n_items, hipStream_t s);
// Plans that move fewer bytes than this run as one launch of every class.
constexpr double kSmallPlanBytes = 16.0 * (1 << 20);
// Tiny-contraction route limits (flops of all core jobs; elements of every payload).
constexpr double kTinyMaxFlops = 16.0 * (1 << 20);
constexpr int64_t kTinyMaxElems = int64_t(1) << 20;
constexpr int kTinyKBlocks = 8;  // jobs with K <= 8 * this load all of K at once
// coeffs[t] = sum over ptr[t] <= p < ptr[t + 1] of scale[p] * src[idx[p]]
// (coefficients of a fused narrow contraction gathered from its small
// operand; t < n).
void launch_coeff_gather(double* coeffs, const double* src, const int64_t* ptr, const int64_t* idx,
                         const double* scale, uint32_t n, hipStream_t s);
// Multi items [begin, begin + n) of at most 4 sources whose coefficients are
// gathered per item (XformArgs::gptr...) into LDS, in one launch with the
// recoupling: the same sums in the same order as launch_coeff_gather, so the
// result equals gather-then-launch_xform bit for bit.  Every block of these
// items must have ndst * nsrc <= kMaxGatherCoef.
constexpr int kMaxGatherCoef = 1024;
void launch_xform_multi_gather(const XformArgs& a, uint32_t begin, uint32_t n, hipStream_t s);

// Device GEMM job: element offsets from the kernel's base pointers.
struct GJob {
  int64_t d_off, a_off, b_off;
  int32_t m, k, n;
  int32_t lda, ldb, ldd;
  int32_t ops;  // bit0: lhs adjoint, bit1: rhs adjoint
  int32_t pad;
};

struct GTile {
  int32_t job;
  int32_t tm;  // tile row index
  int32_t tn;  // tile column index
  int32_t cfg; // tile-shape configuration of this tile's job
};

// Output transform folded into the workgroup-tile epilogue (Single-only
// output transforms whose blocks each map whole core columns): core column
// `col` of a job goes to dst[off + row * rs] scaled by c; off < 0: the
// column is not part of the output.
struct ColMapEntry {
  int64_t off;
  int64_t rs;
  double c;
};

struct GemmArgs {
  const GJob* jobs;
  const GTile* tiles;
  uintptr_t dbase, abase, bbase;  // byte addresses; job offsets are in elements
  double alpha, beta;
  // column-mapped destination (gemm_kernel only; null: D at dbase + d_off in
  // the job's column-major layout): job j's columns at cmap + cmap_start[j]
  const int64_t* cmap_start = nullptr;
  const ColMapEntry* cmap = nullptr;
};

// Gathered small-GEMM route of tiny contractions (one launch per call): the
// operand transforms are folded into the MFMA operand loads and the output
// transform into the epilogue.  One entry per element of a transformed
// operand (value = c * source[idx], c = 0 for elements no block writes) or of
// the core result (stored to dst[idx] scaled by c; idx < 0: stored nowhere).
struct TinyEntry {
  int64_t idx;
  double c;
};

struct TinyArgs {
  const TinyEntry* lmap;  // per element of the transformed lhs (nullptr: lhs used as is)
  const TinyEntry* rmap;  // ... of the transformed rhs
  const TinyEntry* omap;  // per core element (nullptr: the core result is the destination)
  const int64_t* fix;     // destination elements that receive beta * dst (0 when beta == 0)
  uint32_t n_fix;
  uint32_t tile_blocks;   // workgroups of GEMM tiles; the rest sweep `fix`
  double* dst;
};

// gemm_small_kernel's 32 x 32 one-wave tiles with the maps above; tiles of all
// jobs in one launch, FP64.
void launch_gemm_tiny(const GemmArgs& g, const TinyArgs& t, uint32_t n_tiles, hipStream_t s);

// Tile shapes (rows x cols per workgroup) available to the autotuner.
// TILE_SKINNY: narrow outputs (n <= 16), 256 rows x all columns per workgroup
// on the vector units.
// TILE_128x128_W8: 128 x 128 on 8 waves (2 x 4, each 64 x 32), two waves of
// the workgroup per SIMD.
// TILE_SMALL: 32 x 32 per wave, four independent waves per workgroup, operands
// loaded straight into registers (tiny sector GEMMs of batched contractions).
enum TileCfg : int32_t {
  TILE_128x128 = 0,
  TILE_64x64 = 1,
  TILE_128x64 = 2,
  TILE_64x128 = 3,
  TILE_SKINNY = 4,
  TILE_128x128_W8 = 5,
  TILE_SMALL = 6,
  TILE_CFG_COUNT = 7
};
__host__ __device__ inline int tile_rows(int cfg) {
  if (cfg == TILE_SMALL) return 32;
  return cfg == TILE_SKINNY ? 256 : (cfg == TILE_128x128 || cfg == TILE_128x64 || cfg == TILE_128x128_W8) ? 128 : 64;
}
__host__ __device__ inline int tile_cols(int cfg) {
  if (cfg == TILE_SMALL) return 32;
  return cfg == TILE_SKINNY ? 16 : (cfg == TILE_128x128 || cfg == TILE_64x128 || cfg == TILE_128x128_W8) ? 128 : 64;
}

// Launch one grouped GEMM over the tile table.  All tiles of one launch must
// share `cfg` and the op pair `ops` (GJob::ops); the caller splits the table
// by both.
void launch_gemm_f64(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s);
void launch_gemm_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s);
// The same kernels over FP32 payloads (FP64 or FP32 arithmetic).
void launch_gemm_f64_on_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s);
void launch_gemm_f32_on_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s);

// Split-BF16 (Ozaki) GEMM: operands are pre-split into `slices` BF16 planes
// with per-row (lhs) / per-column (rhs) power-of-two scales.
struct SplitArgs {
  const GJob* jobs;
  const GTile* tiles;
  uintptr_t dbase;
  const uint16_t* a_planes;  // [slices][a_elems] bf16, per job packed row-major tiles (see gemm_split.hip)
  const uint16_t* b_planes;
  const int32_t* a_exp;      // per-row exponents
  const int32_t* b_exp;      // per-column exponents
  const int64_t* a_base;     // per job: start of its rows in a_planes / a_exp
  const int64_t* b_base;
  int64_t a_plane_stride, b_plane_stride;
  int slices;
  double alpha, beta;
};

// Reductions / elementwise
void launch_weighted_dot_partials(const double* a, const double* b, const int64_t* chunk_begin,
                                  const int64_t* chunk_end, const int32_t* chunk_region,
                                  const double* weights, double* partials, uint32_t n_chunks,
                                  hipStream_t s);
void launch_sum_partials(const double* partials, uint32_t n, double* out, hipStream_t s);
void launch_axpby(double* y, const double* x, int64_t n, double alpha, double beta, hipStream_t s);
void launch_fill(double* y, int64_t n, double value, hipStream_t s);
void launch_zero_f32(float* y, int64_t n, hipStream_t s);
void launch_axpby_dev(double* y, const double* x, int64_t n, const double* sc, int64_t ai, double af, int64_t bi,
                      double bf, int64_t gi, hipStream_t s);
// A short program of scalar steps on a device scalar buffer (Krylov loops).
enum ScalarOpCode : int32_t {
  SCALAR_SET = 0, SCALAR_DIV = 1, SCALAR_COPY = 2, SCALAR_RSQRT = 3, SCALAR_NEG = 4, SCALAR_CG_GUARD = 5
};
struct ScalarOp {
  int32_t op, out, a, b;
  double imm;
};
constexpr int kMaxScalarOps = 8;
struct ScalarProgram {
  ScalarOp ops[kMaxScalarOps];
  int32_t n;
};
void launch_scalars(double* sc, const ScalarProgram& prog, hipStream_t s);
void launch_fill_unit(double* y, int64_t n, uint64_t seed, hipStream_t s);
void launch_fill_normal(double* y, int64_t n, uint64_t seed, hipStream_t s);

// Complex payloads at the batched GEMM seam (complex.hip): one complex job as
// a real m x 2k x 2n job on planes in a workspace, with an expand pass before
// and a combine pass after the real GEMM.
constexpr int kCplxItem = 1024;  // stored elements per work item of the element passes
enum CplxOps : int32_t { CPLX_A_T = 1, CPLX_A_CONJ = 2, CPLX_B_T = 4, CPLX_B_CONJ = 8 };
struct CplxJob {
  int64_t a_src, b_src, d_dst;       // complex element offsets in lhs, rhs, dst
  int64_t a_ws, b_ws, c_ws;          // double offsets of the job's planes in the workspace
  int32_t m, k, n, ops;              // ops: CplxOps of the stored operands
  uint32_t a_item, b_item, d_item;   // first work item of the job's A / B / D elements
  uint32_t pad;
};
void launch_cplx_expand(const CplxJob* jobs, int n_jobs, uint32_t a_items, uint32_t b_items, const void* lhs,
                        const void* rhs, double* ws, bool c32, hipStream_t s);
void launch_cplx_combine(const CplxJob* jobs, int n_jobs, uint32_t d_items, const double* ws, void* dst,
                         double ar, double ai, double br, double bi, bool c32, hipStream_t s);
// The left-expanded form (C64, op(B) = identity, real beta): R(alpha op(A))
// planes only; the real GEMM reads B and writes D in their real views.
void launch_cplx_expand_lhs(const CplxJob* jobs, int n_jobs, uint32_t a_items, const void* lhs, double* ws,
                            double ar, double ai, hipStream_t s);
void launch_scale_regions(double* y, const int64_t* off, const int64_t* len, uint32_t n,
                          double beta, hipStream_t s);

// Shadow write counts (debug option tnt_ctx_set_shadow): every destination
// element that a work decomposition writes is counted with a 32-bit atomic
// into cnt -- the items of a recoupling plan (each element of each of a
// block's destination layouts), the tiles of a GEMM table (the valid rows and
// columns of each tile), contiguous regions, and the tiny route's stores
// (through its output map) and fix-ups.  launch_shadow_reduce adds the sum of
// cnt[0, n) to stats[0] and the number of elements counted more than once to
// stats[1].
void launch_shadow_plan(const XBlock* blocks, const XLayout* layouts, const XItem* items, uint32_t n_items,
                        uint32_t* cnt, hipStream_t s);
void launch_shadow_tiles(const GJob* jobs, const GTile* tiles, uint32_t n_tiles, uint32_t* cnt, hipStream_t s);
void launch_shadow_regions(const int64_t* off, const int64_t* len, uint32_t n, uint32_t* cnt, hipStream_t s);
void launch_shadow_tiny(const GJob* jobs, const GTile* tiles, uint32_t n_tiles, const TinyEntry* omap,
                        const int64_t* fix, uint32_t n_fix, uint32_t* cnt, hipStream_t s);
void launch_shadow_reduce(const uint32_t* cnt, int64_t n, unsigned long long* stats, hipStream_t s);

}  // namespace tnt_dev
