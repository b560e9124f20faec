// C-ABI of the gfx950 backend (include/tenet_hip.h): contexts, device
// buffers, plan/artifact upload (compiled layouts -> device work lists),
// dispatch of the recoupling and GEMM kernels, reductions, hipGraph capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/tenet_hip.h"
#include "../common/narrow_fusion.hpp"
#include "kernels.hpp"
#include "split.hpp"

using namespace tnt_dev;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct TableCacheEntry {
  void* dev = nullptr;
  size_t bytes = 0;
  uint64_t last_call = 0;  // the entry-point call that last returned it
  int pins = 0;            // live graphs that recorded its address
  std::list<std::string>::iterator lru;  // position in tnt_ctx::table_lru (front = most recent)
};

// Tuned per-shape tile rule (tnt_set_tile_table).
struct TileRule {
  int64_t m_max, n_max;
  int32_t cfg;
};

// The context's GEMM tiling knobs, read when a table is built.
struct TileKnobs {
  int override_cfg = -1;         // tnt_set_tile_config (autotuner sweeps)
  std::vector<TileRule> table;   // tnt_set_tile_table
};

// Entries of the table cache beyond which unpinned entries not used by the
// current call are evicted, least recently used first.
constexpr size_t kTableCacheCap = 512;

struct tnt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::multimap<size_t, void*> free_blocks;  // device pool (size -> ptr)
  std::unordered_map<void*, size_t> live;    // pool-owned allocations
  size_t pool_bytes = 0;
  bool capturing = false;
  // content-addressed device copies of small per-call tables
  std::unordered_map<std::string, TableCacheEntry> tables;
  std::list<std::string> table_lru;
  size_t table_cap = kTableCacheCap;
  uint64_t call = 0;  // incremented by every entry point that uses the cache
  // graphs alive; while any is, memory they may have recorded is retired
  // instead of freed (released when the last graph is destroyed)
  int live_graphs = 0;
  std::vector<void*> retired;
  std::vector<std::string> capture_keys;  // cache entries returned during the open capture
  TileKnobs knobs;
  int split_slices = 5;
  // test hook (tnt_ctx_inject_fault): countdowns to an injected failure, -1 = off
  int fail_alloc_after = -1;
  int fail_launch_after = -1;
  // pinned host scalar for reductions
  double* host_scalar = nullptr;
  double* dev_scalar = nullptr;
  // allocation-free dot (tnt_weighted_dot_async_noalloc): this device's copy
  // of the file-scope `scratch` array, and a host staging area of the same
  // size for its tables, both set up when the context is created
  void* scratch_dev = nullptr;
  std::unique_ptr<unsigned char[]> scratch_host;
  std::string scratch_err;
  std::unique_ptr<SplitWorkspace, SplitWorkspaceDeleter> split;
  // optional per-kernel-class event timing (tnt_ctx_set_timing)
  bool timing = false;
  std::vector<hipEvent_t> event_pool;
  // stream-ordered partials of tnt_weighted_dot_async
  void* dot_parts = nullptr;
  size_t dot_parts_len = 0;
  // real planes of complex GEMM batches (grow-only, stream-ordered)
  void* cplx_ws = nullptr;
  size_t cplx_ws_bytes = 0;
  struct Timed {
    int kind;
    hipEvent_t a, b;
    double work;  // flops (GEMM) or bytes (recoupling)
  };
  std::vector<Timed> timed;
  // debug shadow write counts (tnt_ctx_set_shadow): per-element counts of one
  // destination, and on the device the running totals (writes, elements
  // written more than once); on the host the writes the plans expect
  bool shadow = false;
  uint32_t* shadow_cnt = nullptr;
  size_t shadow_cap = 0;
  unsigned long long* shadow_stats = nullptr;
  uint64_t shadow_expected = 0, shadow_passes = 0;
  // second stream (default priority) for the rhs operand transform of large
  // contractions, forked from and joined back into `stream` by events
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};

static hipEvent_t take_event(tnt_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  // Timing-only events: no system-scope fence at record time (it writes back
  // and invalidates caches, which both slows the following kernel and widens
  // the gap between kernels); stats synchronise the stream before reading.
  hipEvent_t e = nullptr;
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

// Brackets one kernel-class launch with events on the context stream.
// Timing kinds: 0 recoupling replay, 1 + precision one block-GEMM table run
// (all its launches), 8 + 8 * precision + shape one FP64/FP32 GEMM kernel launch.
static int timing_kind_kernel(int precision, int cfg) { return 8 + 8 * precision + cfg; }
constexpr int kTimingKinds = 8 + 8 * 2 + TILE_CFG_COUNT;

struct TimedScope {
  tnt_ctx* ctx;
  int kind;
  double work;
  hipEvent_t a = nullptr;
  TimedScope(tnt_ctx* c, int k, double w) : ctx(c), kind(k), work(w) {
    if (ctx->timing && !ctx->capturing) {
      a = take_event(ctx);
      (void)hipEventRecord(a, ctx->stream);
    }
  }
  ~TimedScope() {
    if (a) {
      hipEvent_t b = take_event(ctx);
      (void)hipEventRecord(b, ctx->stream);
      ctx->timed.push_back({kind, a, b, work});
    }
  }
};

struct tnt_buf {
  void* ptr = nullptr;
  size_t len = 0;
  int dtype = TNT_F64;
  int device = 0;
  bool owned = false;
};

struct DevArrays {  // device copy of a normalised transform plan
  XBlock* blocks = nullptr;
  XLayout* layouts = nullptr;
  XItem* items = nullptr;
  XFlat* flat = nullptr;  // records of the XC_FLAT class, in item order
  double* coeffs = nullptr;
  uint32_t n_items = 0;
  uint32_t cls_begin[XC_COUNT] = {};  // items are grouped by class (one launch each)
  uint32_t cls_count[XC_COUNT] = {};
};

struct tnt_plan {
  int device = 0;
  int64_t dst_len = 0, src_len = 0;
  DevArrays d;
  std::vector<void*> allocs;
  double bytes_moved = 0;  // algorithmic bytes of one replay (read src + write dst)
  bool one_launch = false;  // items of several classes, small: one launch of every class
  uint64_t dst_writes = 0;  // destination element writes of one replay (shadow check)
};

// Plans of several item classes that move fewer bytes than this (default
// kSmallPlanBytes; TENET_AMD_SMALL_PLAN_BYTES overrides at upload) replay in
// one launch.
static double small_plan_bytes() {
  const char* e = std::getenv("TENET_AMD_SMALL_PLAN_BYTES");
  return e ? std::atof(e) : kSmallPlanBytes;
}

struct GemmTable {
  GJob* jobs = nullptr;
  GTile* tiles = nullptr;
  uint32_t n_tiles_cfg[TILE_CFG_COUNT] = {};
  uint32_t tile_start_cfg[TILE_CFG_COUNT] = {};
  // tiles of narrow jobs among each shape's tiles (see part_narrow)
  uint32_t n_narrow_cfg[TILE_CFG_COUNT] = {};
  double flops_cfg[TILE_CFG_COUNT] = {};         // algorithmic flops per shape
  double flops_narrow_cfg[TILE_CFG_COUNT] = {};  // ... of its narrow jobs
  // Within a shape the tiles run in parts of one op pair (GJob::ops), each
  // part its jobs' wide tiles, then its narrow ones: part o of shape c is
  // tiles [part_start[c][o], + part_n[c][o]), the last part_narrow[c][o] of
  // them narrow.
  uint32_t part_start[TILE_CFG_COUNT][4] = {};
  uint32_t part_n[TILE_CFG_COUNT][4] = {};
  uint32_t part_narrow[TILE_CFG_COUNT][4] = {};
  std::vector<GJob> host_jobs;
  std::vector<GTile> host_tiles;
  std::vector<GJob> split_jobs;  // host_jobs with the narrow jobs emptied
  // narrow outputs (n <= 16) on workgroup tiles instead of the narrow-output
  // kernel (tables where those jobs are a small share of the flops)
  bool no_skinny = false;
};

// A job with a dimension of 16 or less is bound by HBM, not by the matrix
// cores: the split-BF16 precision runs it on the exact FP64 kernels instead
// of cutting it into digit planes padded to 16 lines and 32 k.
static bool narrow_job(const GJob& j) { return j.m <= 16 || j.n <= 16 || j.k <= 16; }

struct tnt_artifact {
  int device = 0;
  tnt_plan* lhs = nullptr;
  tnt_plan* rhs = nullptr;
  tnt_plan* out = nullptr;
  GemmTable gemm;
  // output transform folded into the tile GEMM's epilogue (build_colmap):
  // FP64 / FP32 calls on workgroup tiles write the destination directly
  int64_t* cmap_start = nullptr;
  ColMapEntry* cmap = nullptr;
  std::vector<void*> allocs;
  int64_t lhs_len = 0, rhs_len = 0, core_len = 0;
  double* scratch_l = nullptr;
  double* scratch_r = nullptr;
  double* scratch_c = nullptr;
  int64_t* unc_off = nullptr;
  int64_t* unc_len = nullptr;
  uint32_t n_unc = 0;
  int lhs_op = 0, rhs_op = 0;
  // fused narrow route: one recoupling pass lhs -> dst with coefficients
  // gathered from the rhs payload (tnt_fuse::plan_narrow_fusion)
  tnt_plan* fused = nullptr;
  int64_t* g_ptr = nullptr;  // coefficient t: gather terms [g_ptr[t], g_ptr[t + 1])
  int64_t* g_idx = nullptr;
  double* g_scale = nullptr;
  uint32_t n_gather = 0;     // coefficients
  // the fused plan is all few-source Multi items of small blocks: the
  // coefficients are gathered inside its recoupling launch
  bool gather_fold = false;
  // gathered small-GEMM route of tiny contractions (FP64 calls): element maps
  // of the Single-only transforms and one tile list (32 x 32) over every job
  bool tiny = false;
  TinyEntry* t_lmap = nullptr;
  TinyEntry* t_rmap = nullptr;
  TinyEntry* t_omap = nullptr;
  int64_t* t_fix = nullptr;
  uint32_t t_n_fix = 0;
  GJob* t_jobs = nullptr;
  GTile* t_tiles = nullptr;
  uint32_t t_n_tiles = 0;
  double t_flops = 0;
  uint64_t t_writes = 0;     // destination elements the tiny route writes per call
  uint64_t core_writes = 0;  // core elements written by the GEMM jobs and the uncovered regions
};

// The fused narrow route is on unless TENET_AMD_FUSE_NARROW=0 (for A/B runs
// and for tests that compare both routes).
static bool fusion_enabled() {
  const char* e = std::getenv("TENET_AMD_FUSE_NARROW");
  return !(e && e[0] == '0');
}

// Contractions up to this many flops and payload elements whose transforms
// are all Single blocks run as one gathered small-GEMM launch (FP64 calls);
// TENET_AMD_TINY_FLOPS overrides at upload (0: off).  Per call they pay one
// launch instead of one per transform class and tile shape.
static double tiny_max_flops() {
  const char* e = std::getenv("TENET_AMD_TINY_FLOPS");
  return e ? std::atof(e) : kTinyMaxFlops;
}

struct tnt_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<std::string> pinned;  // cache entries this graph recorded
};

#define FAIL(ctx, code, msg)        \
  do {                              \
    if (ctx) (ctx)->err = (msg);    \
    return (code);                  \
  } while (0)

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess) {                                                            \
      if (ctx) (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(_e);         \
      return (_e == hipErrorOutOfMemory) ? TNT_E_OOM : TNT_E_HIP;                      \
    }                                                                                  \
  } while (0)

static int set_device(tnt_ctx* ctx) {
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return TNT_OK;
}

// Status of the launches just queued: the runtime's launch error, or an
// injected one (tnt_ctx_inject_fault).
static int launch_status(tnt_ctx* ctx) {
  if (ctx->fail_launch_after >= 0 && ctx->fail_launch_after-- == 0)
    FAIL(ctx, TNT_E_HIP, "injected launch failure (tnt_ctx_inject_fault)");
  HIPCHK(ctx, hipGetLastError());
  return TNT_OK;
}

// Pooled device allocation: reuse a free block of at least `bytes` and at
// most 2x that; never frees while a graph is being captured.
static int pool_alloc(tnt_ctx* ctx, size_t bytes, void** out) {
  if (ctx->fail_alloc_after >= 0 && ctx->fail_alloc_after-- == 0)
    FAIL(ctx, TNT_E_OOM, "injected allocation failure (tnt_ctx_inject_fault)");
  if (bytes == 0) bytes = 256;
  bytes = (bytes + 255) & ~size_t(255);
  auto it = ctx->free_blocks.lower_bound(bytes);
  if (it != ctx->free_blocks.end() && it->first <= 2 * bytes) {
    *out = it->second;
    ctx->live[it->second] = it->first;
    ctx->free_blocks.erase(it);
    return TNT_OK;
  }
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "device allocation requested during graph capture");
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // release cached blocks and retry once
    for (auto& kv : ctx->free_blocks) {
      (void)hipFree(kv.second);
      ctx->pool_bytes -= kv.first;
    }
    ctx->free_blocks.clear();
    (void)hipGetLastError();
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
      (void)hipGetLastError();
      return TNT_E_OOM;
    }
  }
  ctx->pool_bytes += bytes;
  ctx->live[p] = bytes;
  *out = p;
  return TNT_OK;
}

static void pool_free(tnt_ctx* ctx, void* p) {
  if (!p) return;
  auto it = ctx->live.find(p);
  if (it == ctx->live.end()) return;
  ctx->free_blocks.emplace(it->second, p);
  ctx->live.erase(it);
}

// Release of context-owned memory that a live graph may have recorded: kept
// until the last graph of the context is destroyed.
static void retire(tnt_ctx* ctx, void* p) {
  if (!p) return;
  if (ctx->live_graphs > 0) ctx->retired.push_back(p);
  else pool_free(ctx, p);
}

template <class T>
static int upload_vec(tnt_ctx* ctx, const std::vector<T>& v, T** dev, std::vector<void*>& allocs) {
  void* p = nullptr;
  int rc = pool_alloc(ctx, sizeof(T) * std::max<size_t>(v.size(), 1), &p);
  if (rc) return rc;
  allocs.push_back(p);
  if (!v.empty()) {
    // stream-ordered: a recycled pool block may still be read by queued work
    HIPCHK(ctx, hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  }
  *dev = static_cast<T*>(p);
  return TNT_OK;
}

// Starts an entry-point call that looks tables up in the cache: entries it
// returns are not evicted before the call has queued its kernels.
static void begin_call(tnt_ctx* ctx) { ++ctx->call; }

// Evicts least recently used entries while the cache is over its cap,
// skipping entries pinned by a live graph or returned during the current
// call.  Freed blocks go back to the stream-ordered pool: kernels queued
// earlier on the context stream still read them before any later reuse.
static void trim_tables(tnt_ctx* ctx) {
  auto it = ctx->table_lru.end();
  while (ctx->tables.size() > ctx->table_cap && it != ctx->table_lru.begin()) {
    --it;
    auto e = ctx->tables.find(*it);
    if (e == ctx->tables.end() || e->second.pins > 0 || e->second.last_call == ctx->call) continue;
    pool_free(ctx, e->second.dev);
    ctx->tables.erase(e);
    it = ctx->table_lru.erase(it);
  }
}

// Content-addressed device copy of a small table (job lists, region lists).
static int cached_table(tnt_ctx* ctx, const void* data, size_t bytes, const char* tag, void** dev) {
  std::string key(tag);
  key.append(static_cast<const char*>(data), bytes);
  auto it = ctx->tables.find(key);
  if (it != ctx->tables.end()) {
    TableCacheEntry& e = it->second;
    e.last_call = ctx->call;
    ctx->table_lru.splice(ctx->table_lru.begin(), ctx->table_lru, e.lru);
    if (ctx->capturing) ctx->capture_keys.push_back(key);
    *dev = e.dev;
    return TNT_OK;
  }
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "new job table requested during graph capture; run the call once before capturing");
  void* p = nullptr;
  int rc = pool_alloc(ctx, bytes, &p);
  if (rc) return rc;
  HIPCHK(ctx, hipMemcpyAsync(p, data, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->table_lru.push_front(key);
  TableCacheEntry e;
  e.dev = p;
  e.bytes = bytes;
  e.last_call = ctx->call;
  e.lru = ctx->table_lru.begin();
  ctx->tables.emplace(std::move(key), e);
  trim_tables(ctx);
  *dev = p;
  return TNT_OK;
}

// ---------------------------------------------------------------------------
// Layout normalisation (host, at upload)
// ---------------------------------------------------------------------------
struct NLayout {
  int64_t offset;
  std::vector<int64_t> dims, strides;
};

// Common refinement of the column-major shapes of several layouts with equal
// element counts, then drop extent-1 axes, sort by the first layout's
// strides, and fuse axes contiguous in every layout.
static bool normalize(std::vector<NLayout>& ls, std::vector<int64_t>& dims_out, std::string& err) {
  int64_t count = -1;
  std::vector<int64_t> bps;  // prefix-product breakpoints
  for (auto& l : ls) {
    int64_t p = 1;
    for (auto d : l.dims) {
      if (d <= 0) {
        p = 0;
        break;
      }
      p *= d;
      bps.push_back(p);
    }
    if (count < 0) count = p;
    if (p != count) {
      err = "layouts of one transform block differ in element count";
      return false;
    }
  }
  if (count == 0) {
    dims_out.clear();
    return true;
  }
  std::sort(bps.begin(), bps.end());
  bps.erase(std::unique(bps.begin(), bps.end()), bps.end());
  for (size_t i = 1; i < bps.size(); ++i)
    if (bps[i] % bps[i - 1]) {
      err = "layout shapes have no common refinement";
      return false;
    }
  std::vector<int64_t> ref;  // refined dims
  int64_t prev = 1;
  for (auto b : bps) {
    if (b / prev > 1) ref.push_back(b / prev);
    prev = b;
  }
  // express each layout on the refined dims
  for (auto& l : ls) {
    std::vector<int64_t> st;
    size_t r = 0;
    for (size_t a = 0; a < l.dims.size(); ++a) {
      int64_t rem = l.dims[a], s = l.strides[a];
      while (rem > 1) {
        if (r >= ref.size() || rem % ref[r]) {
          err = "layout refinement failed";
          return false;
        }
        st.push_back(s);
        s *= ref[r];
        rem /= ref[r];
        ++r;
      }
    }
    l.dims = ref;
    l.strides = st;
  }
  size_t n = ref.size();
  std::vector<size_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return ls[0].strides[x] < ls[0].strides[y]; });
  std::vector<int64_t> dims;
  std::vector<std::vector<int64_t>> strides(ls.size());
  for (size_t oi = 0; oi < n; ++oi) {
    size_t a = order[oi];
    bool fuse = !dims.empty();
    for (size_t li = 0; fuse && li < ls.size(); ++li)
      fuse = strides[li].back() * dims.back() == ls[li].strides[a];
    if (fuse) {
      dims.back() *= ref[a];
    } else {
      dims.push_back(ref[a]);
      for (size_t li = 0; li < ls.size(); ++li) strides[li].push_back(ls[li].strides[a]);
    }
  }
  if (dims.empty()) {
    dims.push_back(1);
    for (auto& s : strides) s.push_back(1);
  }
  for (size_t li = 0; li < ls.size(); ++li) ls[li].strides = strides[li];
  dims_out = dims;
  return true;
}

static bool layout_in_bounds(const tnt_layout& l, int64_t len, std::string& err) {
  if (l.rank < 0 || l.rank > TNT_MAX_RANK) {
    err = "layout rank out of range";
    return false;
  }
  int64_t hi = l.offset, n = 1;
  for (int i = 0; i < l.rank; ++i) {
    if (l.shape[i] < 0 || l.strides[i] < 0) {
      err = "negative extent or stride";
      return false;
    }
    n *= l.shape[i];
    if (l.shape[i] > 0) hi += (l.shape[i] - 1) * l.strides[i];
  }
  if (n > 0 && (l.offset < 0 || hi >= len)) {
    err = "layout addresses elements outside its payload";
    return false;
  }
  return true;
}

// Work items per item class that a plan aims for (TENET_AMD_MIN_ITEMS
// overrides at upload; 0 keeps the fixed item sizes).  A small plan cut into
// the fixed sizes leaves most CUs idle: a 190K-element Multi pass is 46 items,
// each swept serially by one workgroup.
static int64_t min_plan_items() {
  const char* e = std::getenv("TENET_AMD_MIN_ITEMS");
  return e ? std::max<int64_t>(0, std::atoll(e)) : kMinPlanItems;
}

// Element updates per Multi work item (TENET_AMD_MULTI_ITEM_WORK overrides at
// upload, for A/B runs).
static int64_t multi_item_work() {
  const char* e = std::getenv("TENET_AMD_MULTI_ITEM_WORK");
  const int64_t v = e ? std::atoll(e) : 0;
  return v >= 256 ? v : kMultiItemWork;
}

static int build_plan(tnt_ctx* ctx, const tnt_transform_desc* d, tnt_plan** out) {
  if (!d) FAIL(ctx, TNT_E_INVALID, "null transform descriptor");
  std::vector<XBlock> blocks;
  std::vector<XLayout> layouts;
  std::vector<XItem> cls_items[XC_COUNT];
  std::vector<XFlat> flat;
  std::string err;
  double bytes = 0;
  uint64_t dst_writes = 0;
  // item sizes: Single kSingleItem elements, Multi ~kMultiItemWork element
  // updates (elements x layouts / 2), both shrunk for small plans so that each
  // class has about min_plan_items() items (never below 256 elements)
  int64_t single_target = kSingleItem, multi_work_target = multi_item_work();
  {
    int64_t single_elems = 0, multi_work = 0;
    for (size_t bi = 0; bi < d->n_blocks; ++bi) {
      const tnt_transform_block& tb = d->blocks[bi];
      if (tb.dst_layout >= d->n_layouts) continue;  // rejected below
      const tnt_layout& L = d->layouts[tb.dst_layout];
      if (L.rank < 0 || L.rank > TNT_MAX_RANK) continue;
      int64_t n = 1;
      for (int i = 0; i < L.rank; ++i) n *= std::max<int64_t>(0, L.shape[i]);
      if (tb.kind == TNT_BLOCK_SINGLE) single_elems += n;
      else multi_work += n * std::max<int64_t>(1, ((int64_t)tb.dst_count + tb.src_count) / 2);
    }
    const int64_t want = min_plan_items();
    if (want > 0 && single_elems < want * kSingleItem)
      single_target = std::max<int64_t>(256, ((single_elems / want + 63) / 64) * 64);
    if (want > 0 && multi_work < want * multi_work_target)
      multi_work_target = std::max<int64_t>(256, multi_work / want);
  }
  auto add_block = [&](int kind, const std::vector<uint32_t>& dl, const std::vector<uint32_t>& sl,
                       uint32_t coeff) -> int {
    std::vector<NLayout> ls;
    for (auto i : dl) {
      if (i >= d->n_layouts) FAIL(ctx, TNT_E_OOB, "layout index out of range");
      const tnt_layout& L = d->layouts[i];
      if (!layout_in_bounds(L, d->dst_len, err)) FAIL(ctx, TNT_E_OOB, "destination " + err);
      ls.push_back({L.offset, std::vector<int64_t>(L.shape, L.shape + L.rank),
                    std::vector<int64_t>(L.strides, L.strides + L.rank)});
    }
    for (auto i : sl) {
      if (i >= d->n_layouts) FAIL(ctx, TNT_E_OOB, "layout index out of range");
      const tnt_layout& L = d->layouts[i];
      if (!layout_in_bounds(L, d->src_len, err)) FAIL(ctx, TNT_E_OOB, "source " + err);
      ls.push_back({L.offset, std::vector<int64_t>(L.shape, L.shape + L.rank),
                    std::vector<int64_t>(L.strides, L.strides + L.rank)});
    }
    std::vector<int64_t> dims;
    if (!normalize(ls, dims, err)) FAIL(ctx, TNT_E_SHAPE, err);
    if (dims.empty()) return TNT_OK;  // zero-size block
    if (dims.size() > (size_t)kMaxRank) FAIL(ctx, TNT_E_UNSUPPORTED, "normalised rank exceeds 8");
    // A Single block contiguous along another axis in its source than in its
    // destination (dims[0], smallest destination stride) is recoupled through
    // LDS tiles: that axis becomes dims[1].
    if (kind == XK_SINGLE && dims.size() >= 2 && ls[0].strides[0] == 1 && ls[1].strides[0] != 1 &&
        dims[0] >= 16) {
      for (size_t t = 1; t < dims.size(); ++t) {
        if (ls[1].strides[t] == 1 && dims[t] >= 16) {
          std::swap(dims[1], dims[t]);
          for (auto& l : ls) std::swap(l.strides[1], l.strides[t]);
          kind = XK_TRANSPOSE;
          break;
        }
      }
    }
    XBlock b{};
    b.kind = kind;
    b.rank = (int32_t)dims.size();
    b.ndst = (int32_t)dl.size();
    b.nsrc = (int32_t)sl.size();
    b.lay = (int32_t)layouts.size();
    b.coeff = (int32_t)coeff;
    for (size_t i = 0; i < dims.size(); ++i) b.dims[i] = dims[i];
    for (auto& l : ls) {
      XLayout x{};
      x.offset = l.offset;
      for (size_t i = 0; i < l.strides.size(); ++i) x.strides[i] = l.strides[i];
      layouts.push_back(x);
    }
    uint32_t bi = (uint32_t)blocks.size();
    blocks.push_back(b);
    int64_t plane = dims[0] * (dims.size() > 1 ? dims[1] : 1);
    int64_t outer = 1;
    for (size_t i = 2; i < dims.size(); ++i) outer *= dims[i];
    int64_t elems = plane * outer;
    bytes += 8.0 * double(elems) * double(dl.size() + sl.size());
    dst_writes += (uint64_t)elems * dl.size();
    // Single blocks: items of up to kSingleItem elements (kSingleUnroll per
    // thread) swept in order by a persistent grid; Multi: ~kMultiItemWork
    // element-updates per item, weighted by its layouts (both smaller for small
    // plans).  Single items of one outer index go to the flattened class.
    struct ItemList {
      std::vector<XItem>& items;
      std::vector<XItem>& flat_items;
      std::vector<XFlat>& flat;
      const XBlock& b;
      const std::vector<NLayout>& ls;
      double c;
      void push_back(const XItem& it) {
        if (b.kind != XK_SINGLE || it.o_count != 1) {
          items.push_back(it);
          return;
        }
        int64_t od = ls[0].offset, os = ls[1].offset;
        uint64_t o = it.o_begin;
        for (int d = 2; d < b.rank; ++d) {
          const int64_t i = (int64_t)(o % (uint64_t)b.dims[d]);
          o /= (uint64_t)b.dims[d];
          od += i * ls[0].strides[d];
          os += i * ls[1].strides[d];
        }
        XFlat f{};
        f.od = od;
        f.os = os;
        f.d0 = ls[0].strides[0];
        f.s0 = ls[1].strides[0];
        f.d1 = b.rank > 1 ? ls[0].strides[1] : 0;
        f.s1 = b.rank > 1 ? ls[1].strides[1] : 0;
        f.c = c;
        f.n0 = (uint32_t)b.dims[0];
        f.e_begin = it.e_begin;
        f.e_count = it.e_count;
        flat_items.push_back(it);
        flat.push_back(f);
      }
    } items{cls_items[xclass_of(kind, (int)sl.size())], cls_items[XC_FLAT], flat, blocks.back(), ls,
            kind == XK_SINGLE ? d->coeffs[coeff] : 0.0};
    if (kind == XK_TRANSPOSE) {  // one item per 64 x 64 tile of every plane
      const int64_t ntx = (dims[0] + kTransposeTile - 1) / kTransposeTile;
      const int64_t nty = (dims[1] + kTransposeTile - 1) / kTransposeTile;
      for (int64_t o = 0; o < outer; ++o)
        for (int64_t t = 0; t < ntx * nty; ++t)
          items.push_back({bi, 1u, (uint64_t)o, (uint32_t)t, (uint32_t)kTransposeTile});
      return TNT_OK;
    }
    int64_t weight = std::max<int64_t>(1, (int64_t)(dl.size() + sl.size()) / 2);
    int64_t target = kind == XK_SINGLE ? single_target : std::max<int64_t>(256, multi_work_target / weight);
    if (plane >= target) {
      int64_t chunk = ((target + 63) / 64) * 64;
      for (int64_t o = 0; o < outer; ++o)
        for (int64_t e = 0; e < plane; e += chunk)
          items.push_back({bi, 1u, (uint64_t)o, (uint32_t)e, (uint32_t)std::min(chunk, plane - e)});
    } else {
      int64_t per = std::max<int64_t>(1, target / plane);
      for (int64_t o = 0; o < outer; o += per)
        items.push_back({bi, (uint32_t)std::min(per, outer - o), (uint64_t)o, 0u, (uint32_t)plane});
    }
    return TNT_OK;
  };
  for (size_t bi = 0; bi < d->n_blocks; ++bi) {
    const tnt_transform_block& tb = d->blocks[bi];
    std::vector<uint32_t> dl, sl;
    for (uint32_t i = 0; i < tb.dst_count; ++i) dl.push_back(tb.dst_layout + i);
    for (uint32_t i = 0; i < tb.src_count; ++i) sl.push_back(tb.src_layout + i);
    int rc;
    if (tb.kind == TNT_BLOCK_SINGLE) {
      if (tb.dst_count != 1 || tb.src_count != 1) FAIL(ctx, TNT_E_INVALID, "Single block needs one source and one destination");
      if (tb.coeff >= d->n_coeffs) FAIL(ctx, TNT_E_OOB, "coefficient index out of range");
      rc = add_block(XK_SINGLE, dl, sl, tb.coeff);
    } else if (tb.kind == TNT_BLOCK_MULTI) {
      if (tb.src_count == 0 || tb.dst_count == 0) FAIL(ctx, TNT_E_INVALID, "empty Multi block");
      if (tb.src_count > (uint32_t)kMaxMultiSrc)
        FAIL(ctx, TNT_E_UNSUPPORTED, "Multi block with more than 32 source blocks");
      if ((size_t)tb.coeff + (size_t)tb.src_count * tb.dst_count > d->n_coeffs)
        FAIL(ctx, TNT_E_OOB, "coefficient range out of bounds");
      rc = add_block(XK_MULTI, dl, sl, tb.coeff);
    } else {
      FAIL(ctx, TNT_E_INVALID, "unknown transform block kind");
    }
    if (rc) return rc;
  }
  for (size_t i = 0; i < d->n_inactive; ++i) {
    int rc = add_block(XK_SCALE, {d->inactive[i]}, {}, 0);
    if (rc) return rc;
  }
  auto* p = new tnt_plan;
  p->device = ctx->device;
  p->dst_len = d->dst_len;
  p->src_len = d->src_len;
  p->bytes_moved = bytes;
  p->dst_writes = dst_writes;
  std::vector<double> coeffs(d->coeffs, d->coeffs + d->n_coeffs);
  std::vector<XItem> items;
  for (int c = 0; c < XC_COUNT; ++c) {
    p->d.cls_begin[c] = (uint32_t)items.size();
    p->d.cls_count[c] = (uint32_t)cls_items[c].size();
    items.insert(items.end(), cls_items[c].begin(), cls_items[c].end());
  }
  int rc = upload_vec(ctx, blocks, &p->d.blocks, p->allocs);
  if (!rc) rc = upload_vec(ctx, layouts, &p->d.layouts, p->allocs);
  if (!rc) rc = upload_vec(ctx, items, &p->d.items, p->allocs);
  if (!rc && !flat.empty()) rc = upload_vec(ctx, flat, &p->d.flat, p->allocs);
  if (!rc) rc = upload_vec(ctx, coeffs, &p->d.coeffs, p->allocs);
  if (rc) {
    for (auto a : p->allocs) pool_free(ctx, a);
    delete p;
    return rc;
  }
  p->d.n_items = (uint32_t)items.size();
  int classes = 0;  // launches of a replay (flattened and other Single items share one)
  for (int c = 0; c < XC_COUNT; ++c)
    if (c != XC_FLAT) classes += p->d.cls_count[c] > 0 || (c == XC_SINGLE && p->d.cls_count[XC_FLAT] > 0);
  p->one_launch = classes > 1 && bytes < small_plan_bytes();
  *out = p;
  return TNT_OK;
}

// Shadow check of one destination (debug option): counts the writes of a
// work decomposition (launch(cnt)) over `len` elements and adds them to the
// context's totals, with the writes the decomposition is planned to make.
// Not while capturing a graph.  Returns non-zero only when the count buffer
// cannot be allocated.
extern "C++" template <class F>
static int shadow_pass(tnt_ctx* ctx, int64_t len, uint64_t expected, F&& launch) {
  if (!ctx->shadow || ctx->capturing || len <= 0) return TNT_OK;
  if ((size_t)len > ctx->shadow_cap) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    (void)hipFree(ctx->shadow_cnt);
    ctx->shadow_cnt = nullptr;
    ctx->shadow_cap = 0;
    HIPCHK(ctx, hipMalloc(&ctx->shadow_cnt, sizeof(uint32_t) * (size_t)len));
    ctx->shadow_cap = (size_t)len;
  }
  launch_zero_f32(reinterpret_cast<float*>(ctx->shadow_cnt), len, ctx->stream);
  launch(ctx->shadow_cnt);
  launch_shadow_reduce(ctx->shadow_cnt, len, ctx->shadow_stats, ctx->stream);
  ctx->shadow_expected += expected;
  ctx->shadow_passes += 1;
  return TNT_OK;
}

static void run_plan(tnt_ctx* ctx, const tnt_plan* p, double* dst, const double* src, double alpha,
                     double beta, int overwrite) {
  XformArgs a{p->d.blocks, p->d.layouts, p->d.items, p->d.flat, p->d.coeffs, dst, src, alpha, beta, overwrite};
  {
    TimedScope ts(ctx, 0, p->bytes_moved);
    if (p->one_launch) {
      launch_xform_any(a, p->d.n_items, ctx->stream);
    } else {
      launch_xform_single(a, p->d.cls_begin[XC_SINGLE], p->d.cls_count[XC_SINGLE], p->d.cls_count[XC_FLAT],
                          ctx->stream);
      for (int c = 0; c < XC_COUNT; ++c)
        if (c != XC_SINGLE && c != XC_FLAT) launch_xform(a, c, p->d.cls_begin[c], p->d.cls_count[c], ctx->stream);
    }
  }
  (void)shadow_pass(ctx, p->dst_len, p->dst_writes, [&](uint32_t* cnt) {
    launch_shadow_plan(p->d.blocks, p->d.layouts, p->d.items, p->d.n_items, cnt, ctx->stream);
  });
}

// ---------------------------------------------------------------------------
// GEMM tables + tile-shape autotuner
// ---------------------------------------------------------------------------
// Tile shape per job: the context's measured preference table (installed by
// the autotuner), else the size rule below.
// Table-level default: the largest tile shape that still gives every CU at
// least ~2 (128x128) or 1 (64x128 / 128x64) workgroups and covers the jobs
// with little padding (>= 80 % of its area used), counted over the jobs that
// take a workgroup tile; narrow outputs go to the vector-unit kernel.  The
// one-wave 32 x 32 tiles (TILE_SMALL) are chosen by measurement only (an
// extra tile shape is an extra launch).  A per-artifact override (autotuner)
// wins.
static int default_family(const std::vector<GJob>& jobs) {
  auto count = [&](int cfg, double* used_frac) {
    int64_t t = 0;
    double used = 0, padded = 0;
    const int bm = tile_rows(cfg), bn = tile_cols(cfg);
    for (const auto& j : jobs) {
      if (j.m <= 0 || j.n <= 0 || j.n <= 16) continue;
      const int64_t tm = (j.m + bm - 1) / bm, tn = (j.n + bn - 1) / bn;
      t += tm * tn;
      used += double(j.m) * double(j.n);
      padded += double(tm * bm) * double(tn * bn);
    }
    *used_frac = padded > 0 ? used / padded : 0.0;
    return t;
  };
  // 128 x 128 tiles where they fill the chip four times over with little
  // padding (large sectors: 82 -> 85 % of peak on 4096^3 against 64 x 128);
  // otherwise 64 x 64 -- on uneven small-K sector GEMMs (the U(1) environment
  // update) as fast as 64 x 128 or faster (L.A: 65 % against 59.5 % of peak),
  // and one shape per contraction is one launch.
  double f128 = 0;
  if (count(TILE_128x128, &f128) >= 4 * 256 && f128 >= 0.8) return TILE_128x128;
  return TILE_64x64;
}

static int job_cfg(const GJob& j, int family, const TileKnobs& kn) {
  if (j.n <= 16) return TILE_SKINNY;
  for (const auto& r : kn.table)
    if (j.m <= r.m_max && j.n <= r.n_max) return r.cfg;
  int cfg = family;
  // keep narrow dimensions off oversized tiles
  if (cfg == TILE_128x128 && j.m < 128 && j.n < 128) cfg = TILE_64x64;
  else if (cfg == TILE_128x128 && j.m < 128) cfg = TILE_64x128;
  else if (cfg == TILE_128x128 && j.n < 128) cfg = TILE_128x64;
  else if (cfg == TILE_64x128 && j.m > 2 * j.n) cfg = TILE_128x64;
  return cfg;
}

static void build_gemm_table(GemmTable& t, const std::vector<GJob>& jobs, const TileKnobs& kn, int forced = -1,
                             bool no_skinny = false) {
  t.host_jobs = jobs;
  t.no_skinny = no_skinny;
  std::vector<std::vector<GTile>> per(TILE_CFG_COUNT);
  // larger jobs first for load balance
  std::vector<size_t> order(jobs.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    double ca = double(jobs[a].m) * jobs[a].n * jobs[a].k, cb = double(jobs[b].m) * jobs[b].n * jobs[b].k;
    return ca > cb;
  });
  const int ovr = forced >= 0 ? forced : kn.override_cfg;
  const int family = default_family(jobs);
  for (size_t oi : order) {
    const GJob& j = jobs[oi];
    if (j.m <= 0 || j.n <= 0) continue;
    int cfg = job_cfg(j, family, kn);
    // a forced shape leaves narrow outputs on the vector kernel, except that
    // forced one-wave tiles also take narrow jobs of few rows (one launch for
    // a batch of tiny GEMMs instead of two)
    if (ovr >= 0 && (cfg != TILE_SKINNY || (ovr == TILE_SMALL && j.m <= 256))) cfg = ovr;
    if (cfg == TILE_SKINNY && j.n > 16) cfg = TILE_64x64;
    if (cfg == TILE_SKINNY && no_skinny) cfg = ovr >= 0 && ovr != TILE_SKINNY && ovr != TILE_SMALL ? ovr : TILE_64x64;
    // nothing contracted: D = beta D, written by the workgroup-tile epilogue
    // (that kernel loads no operand element when k == 0), on the table's
    // workgroup shape so that such jobs join its launch
    if (j.k == 0) cfg = ovr >= 0 && ovr != TILE_SKINNY && ovr != TILE_SMALL ? ovr : family;
    int bm = tile_rows(cfg), bn = tile_cols(cfg);
    int tm = (j.m + bm - 1) / bm, tn = (j.n + bn - 1) / bn;
    for (int a = 0; a < tm; ++a)
      for (int b = 0; b < tn; ++b) per[cfg].push_back({(int32_t)oi, a, b, cfg});
  }
  t.host_tiles.clear();
  for (int c = 0; c < TILE_CFG_COUNT; ++c) {
    t.flops_cfg[c] = t.flops_narrow_cfg[c] = 0.0;
    int32_t last = -1;
    for (const GTile& g : per[c]) {  // tiles of one job are adjacent
      if (g.job == last) continue;
      last = g.job;
      const GJob& j = jobs[g.job];
      const double f = 2.0 * double(j.m) * double(j.n) * double(j.k);
      t.flops_cfg[c] += f;
      if (narrow_job(j)) t.flops_narrow_cfg[c] += f;
    }
    // parts by op pair; in each, narrow jobs' tiles last (stable: larger
    // jobs still lead)
    t.tile_start_cfg[c] = (uint32_t)t.host_tiles.size();
    t.n_tiles_cfg[c] = (uint32_t)per[c].size();
    t.n_narrow_cfg[c] = 0;
    for (int o = 0; o < 4; ++o) {
      std::vector<GTile> part;
      for (const GTile& g : per[c])
        if ((jobs[g.job].ops & 3) == o) part.push_back(g);
      auto mid = std::stable_partition(part.begin(), part.end(),
                                       [&](const GTile& g) { return !narrow_job(jobs[g.job]); });
      t.part_start[c][o] = (uint32_t)t.host_tiles.size();
      t.part_n[c][o] = (uint32_t)part.size();
      t.part_narrow[c][o] = (uint32_t)(part.end() - mid);
      t.n_narrow_cfg[c] += t.part_narrow[c][o];
      t.host_tiles.insert(t.host_tiles.end(), part.begin(), part.end());
    }
  }
  t.split_jobs = jobs;
  for (auto& j : t.split_jobs)
    if (narrow_job(j)) j.m = j.n = 0;
}

static int upload_gemm_table(tnt_ctx* ctx, GemmTable& t, std::vector<void*>& allocs) {
  int rc = upload_vec(ctx, t.host_jobs, &t.jobs, allocs);
  if (!rc) rc = upload_vec(ctx, t.host_tiles, &t.tiles, allocs);
  return rc;
}

static double table_flops(const GemmTable& t) {
  double f = 0;
  for (const auto& j : t.host_jobs) f += 2.0 * double(j.m) * double(j.n) * double(j.k);
  return f;
}

// storage: payload type of d, a and b (TNT_F64, or TNT_F32 for FP64 / FP32
// arithmetic only; the caller has checked the combination).
// side / side_done: operand side_op (0 lhs, 1 rhs) is still being written on
// `side` (its transform was forked there); `side_done` is recorded there
// after that work.  The split-INT8 table runs that operand's pre-passes on
// `side` and the other's here; every other table waits for `side_done` first.
static int run_gemm_table(tnt_ctx* ctx, const GemmTable& t, uintptr_t d, uintptr_t a, uintptr_t b,
                          double alpha, double beta, int precision, int storage = TNT_F64,
                          const int64_t* cmap_start = nullptr, const ColMapEntry* cmap = nullptr,
                          hipStream_t side = nullptr, hipEvent_t side_done = nullptr, int side_op = 1) {
  if (side && precision != TNT_PREC_SPLIT_I8) {
    HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, side_done, 0));
    side = nullptr;
  }
  TimedScope ts(ctx, 1 + precision, ctx->timing ? table_flops(t) : 0.0);
  if (storage == TNT_F32) {
    for (int c = 0; c < TILE_CFG_COUNT; ++c) {
      if (!t.n_tiles_cfg[c]) continue;
      TimedScope tk(ctx, timing_kind_kernel(precision, c), ctx->timing ? t.flops_cfg[c] : 0.0);
      for (int o = 0; o < 4; ++o) {
        if (!t.part_n[c][o]) continue;
        GemmArgs g{t.jobs, t.tiles + t.part_start[c][o], d, a, b, alpha, beta};
        if (precision == TNT_PREC_F32) launch_gemm_f32_on_f32(g, c, o, t.part_n[c][o], ctx->stream);
        else launch_gemm_f64_on_f32(g, c, o, t.part_n[c][o], ctx->stream);
      }
    }
    return launch_status(ctx);
  }
  if (precision == TNT_PREC_SPLIT_BF16 || precision == TNT_PREC_SPLIT_I8) {
    // FP64 kernels for the narrow jobs (after the split run when the rhs is
    // still being written on `side`: that run joins it)
    auto narrow = [&]() {
      for (int c = 0; c < TILE_CFG_COUNT; ++c) {
        if (!t.n_narrow_cfg[c]) continue;
        TimedScope tk(ctx, timing_kind_kernel(TNT_PREC_F64, c), ctx->timing ? t.flops_narrow_cfg[c] : 0.0);
        for (int o = 0; o < 4; ++o) {
          if (!t.part_narrow[c][o]) continue;
          const uint32_t first = t.part_start[c][o] + t.part_n[c][o] - t.part_narrow[c][o];
          GemmArgs g{t.jobs, t.tiles + first, d, a, b, alpha, beta};
          launch_gemm_f64(g, c, o, t.part_narrow[c][o], ctx->stream);
        }
      }
    };
    if (!side) narrow();
    if (!ctx->split) ctx->split.reset(split_workspace_create());
    const int scheme = precision == TNT_PREC_SPLIT_I8 ? SPLIT_SCHEME_I8 : SPLIT_SCHEME_BF16;
    int rc = split_gemm_run(ctx->split.get(), scheme, ctx->split_slices, t.split_jobs, t.jobs, d, a, b, alpha, beta,
                            ctx->stream, ctx->capturing, ctx->live_graphs > 0, ctx->err, side, side_done, side_op);
    if (side) {
      // joined here too: the run may have returned before joining
      HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, side_done, 0));
      if (!rc) narrow();
    }
    if (rc) return rc;
    return launch_status(ctx);
  }
  for (int c = 0; c < TILE_CFG_COUNT; ++c) {
    if (!t.n_tiles_cfg[c]) continue;
    TimedScope tk(ctx, timing_kind_kernel(precision, c), ctx->timing ? t.flops_cfg[c] : 0.0);
    for (int o = 0; o < 4; ++o) {
      if (!t.part_n[c][o]) continue;
      GemmArgs g{t.jobs, t.tiles + t.part_start[c][o], d, a, b, alpha, beta, cmap_start, cmap};
      if (precision == TNT_PREC_F32) launch_gemm_f32(g, c, o, t.part_n[c][o], ctx->stream);
      else launch_gemm_f64(g, c, o, t.part_n[c][o], ctx->stream);
    }
  }
  return launch_status(ctx);
}

static bool job_in_bounds(const GJob& j, size_t dlen, size_t alen, size_t blen) {
  auto span = [](int64_t off, int64_t rows, int64_t cols, int64_t ld) -> int64_t {
    if (rows == 0 || cols == 0) return off;
    return off + (cols - 1) * ld + rows;  // one past last element
  };
  bool a_adj = j.ops & 1, b_adj = j.ops & 2;
  int64_t ar = a_adj ? j.k : j.m, ac = a_adj ? j.m : j.k;
  int64_t br = b_adj ? j.n : j.k, bc = b_adj ? j.k : j.n;
  if (j.lda < std::max<int64_t>(1, ar) || j.ldb < std::max<int64_t>(1, br) || j.ldd < std::max<int64_t>(1, j.m))
    return false;
  if (j.d_off < 0 || j.a_off < 0 || j.b_off < 0) return false;
  return span(j.d_off, j.m, j.n, j.ldd) <= (int64_t)dlen && span(j.a_off, ar, ac, j.lda) <= (int64_t)alen &&
         span(j.b_off, br, bc, j.ldb) <= (int64_t)blen;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free dot (tnt_weighted_dot_async_noalloc)
// ---------------------------------------------------------------------------
// Chunk tables, weights and per-chunk partials of one call:
//   [weights: n_regions f64][begins: n_chunks i64][ends: n_chunks i64]
//   [region ids: n_chunks i32, padded to 8 bytes][partials: n_chunks f64]
// Each GPU holds its own copy of the array until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: every
// host use goes through the address hipGetSymbolAddress returned for the
// device (tnt_ctx::scratch_dev).
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
static void* g_scratch_dev[kMaxScratchDevices];
static std::mutex g_scratch_mu;

// Resolves this device's address of `scratch` once per device (the device is
// current: called from tnt_ctx_create).  A failure is kept in scratch_err and
// reported by the allocation-free entry point only.
static void scratch_setup(tnt_ctx* ctx) {
  ctx->scratch_host.reset(new unsigned char[kNoAllocScratchBytes]);
  if (ctx->device < 0 || ctx->device >= kMaxScratchDevices) {
    ctx->scratch_err = "device ordinal beyond the allocation-free scratch table";
    return;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      ctx->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    g_scratch_dev[ctx->device] = ptr;
  }
  ctx->scratch_dev = g_scratch_dev[ctx->device];
}

// MatrixOp values (Identity, Transpose, Adjoint); on real payloads both
// non-identity ops transpose the stored matrix.
static bool valid_op(int op) { return op == TNT_OP_IDENTITY || op == TNT_OP_TRANSPOSE || op == TNT_OP_ADJOINT; }
static bool op_transposes(int op) { return op == TNT_OP_TRANSPOSE || op == TNT_OP_ADJOINT; }

static bool valid_precision(int p) { return p >= TNT_PREC_F64 && p <= TNT_PREC_SPLIT_I8; }

// Payload type check of tnt_buf_alloc / tnt_buf_wrap.
static int check_dtype(tnt_ctx* ctx, int dtype) {
  if (dtype == TNT_F64 || dtype == TNT_F32 || dtype == TNT_C32 || dtype == TNT_C64) return TNT_OK;
  FAIL(ctx, TNT_E_INVALID, "unknown payload type");
}

static bool is_complex(int dtype) { return dtype == TNT_C32 || dtype == TNT_C64; }
static size_t dtype_bytes(int dtype) { return dtype == TNT_F32 ? 4 : dtype == TNT_C64 ? 16 : 8; }

// The job of a DenseGemmBatchJob against whole flat payloads (ld = row count
// of the stored operand).  Dimensions must fit the kernels' 32-bit indices.
static int make_seam_job(tnt_ctx* ctx, const tnt_gemm_job& j, int lhs_op, int rhs_op, GJob* out) {
  if (j.rows > INT32_MAX || j.cols > INT32_MAX || j.contracted > INT32_MAX)
    FAIL(ctx, TNT_E_UNSUPPORTED, "job dimension exceeds 2^31 - 1");
  if (j.dst_offset > (size_t)INT64_MAX || j.lhs_offset > (size_t)INT64_MAX || j.rhs_offset > (size_t)INT64_MAX)
    FAIL(ctx, TNT_E_OOB, "job offset exceeds 2^63 - 1");
  GJob g{};
  g.d_off = (int64_t)j.dst_offset;
  g.a_off = (int64_t)j.lhs_offset;
  g.b_off = (int64_t)j.rhs_offset;
  g.m = (int32_t)j.rows;
  g.k = (int32_t)j.contracted;
  g.n = (int32_t)j.cols;
  const bool la = op_transposes(lhs_op), rb = op_transposes(rhs_op);
  g.lda = la ? std::max(g.k, 1) : std::max(g.m, 1);
  g.ldb = rb ? std::max(g.n, 1) : std::max(g.k, 1);
  g.ldd = std::max(g.m, 1);
  g.ops = (la ? 1 : 0) | (rb ? 2 : 0);
  *out = g;
  return TNT_OK;
}

// Core regions that no job writes as GEMM jobs with nothing contracted (the
// workgroup-tile epilogue gives dst = beta dst): region [off, off + len) is
// a 64-row column-major block of len / 64 columns (at most 2^30 columns per
// job) plus one column of len % 64 rows.  ops (the contraction's op pair)
// keeps them in the launch of its other tiles; no operand element is read.
static void region_jobs(const tnt_region* r, size_t n, int32_t ops, std::vector<GJob>& out) {
  constexpr int64_t kRows = 64, kMaxCols = int64_t(1) << 30;
  auto push = [&](int64_t off, int64_t m, int64_t cols) {
    GJob g{};
    g.d_off = off;
    g.m = (int32_t)m;
    g.n = (int32_t)cols;
    g.k = 0;
    g.lda = (int32_t)m;
    g.ldb = (int32_t)cols;
    g.ldd = (int32_t)m;
    g.ops = ops;
    out.push_back(g);
  };
  for (size_t i = 0; i < n; ++i) {
    int64_t off = r[i].offset, cols = r[i].len / kRows;
    while (cols > 0) {
      const int64_t c = std::min(cols, kMaxCols);
      push(off, kRows, c);
      off += kRows * c;
      cols -= c;
    }
    if (r[i].len % kRows) push(off, r[i].len % kRows, 1);
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int tnt_ctx_create(int device, tnt_ctx** out) {
  if (!out) return TNT_E_INVALID;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) return TNT_E_HIP;
  if (device < 0 || device >= n) return TNT_E_DEVICE_MISMATCH;
  auto* c = new tnt_ctx;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->host_scalar), sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&c->dev_scalar), sizeof(double)) != hipSuccess) {
    delete c;
    return TNT_E_HIP;
  }
  scratch_setup(c);
  *out = c;
  return TNT_OK;
}

int tnt_ctx_destroy(tnt_ctx* ctx) {
  if (!ctx) return TNT_E_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->split.reset();
  for (auto p : ctx->retired) pool_free(ctx, p);
  ctx->retired.clear();
  for (auto& kv : ctx->tables) pool_free(ctx, kv.second.dev);
  for (auto& kv : ctx->live) (void)hipFree(kv.first);
  for (auto& kv : ctx->free_blocks) (void)hipFree(kv.second);
  for (auto& t : ctx->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  (void)hipFree(ctx->dev_scalar);
  (void)hipHostFree(ctx->host_scalar);
  (void)hipFree(ctx->shadow_cnt);
  (void)hipFree(ctx->shadow_stats);
  if (ctx->side) {
    (void)hipStreamSynchronize(ctx->side);
    (void)hipStreamDestroy(ctx->side);
  }
  if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
  if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return TNT_OK;
}

const char* tnt_last_error(const tnt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tnt_ctx_synchronize(tnt_ctx* ctx) {
  if (!ctx) return TNT_E_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TNT_OK;
}

int tnt_ctx_stream(tnt_ctx* ctx, void** stream) {
  if (!ctx || !stream) return TNT_E_INVALID;
  *stream = ctx->stream;
  return TNT_OK;
}

int tnt_ctx_pool_bytes(const tnt_ctx* ctx, size_t* bytes) {
  if (!ctx || !bytes) return TNT_E_INVALID;
  *bytes = ctx->pool_bytes;
  return TNT_OK;
}

int tnt_buf_alloc(tnt_ctx* ctx, size_t len, int dtype, tnt_buf** out) {
  if (!ctx || !out) return TNT_E_INVALID;
  if (int rc = check_dtype(ctx, dtype)) return rc;
  if (len > (SIZE_MAX >> 4)) FAIL(ctx, TNT_E_OOM, "payload length too large");
  if (int rc = set_device(ctx)) return rc;
  auto* b = new tnt_buf;
  b->len = len;
  b->dtype = dtype;
  b->device = ctx->device;
  b->owned = true;
  int rc = pool_alloc(ctx, len * dtype_bytes(dtype), &b->ptr);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return TNT_OK;
}

int tnt_buf_wrap(tnt_ctx* ctx, void* ptr, size_t len, int dtype, tnt_buf** out) {
  if (!ctx || !out || (!ptr && len)) return TNT_E_INVALID;
  if (int rc = check_dtype(ctx, dtype)) return rc;
  hipPointerAttribute_t attr;
  if (ptr && hipPointerGetAttributes(&attr, ptr) == hipSuccess) {
    if (attr.type == hipMemoryTypeDevice && attr.device != ctx->device)
      FAIL(ctx, TNT_E_DEVICE_MISMATCH, "wrapped pointer lives on another device");
  }
  (void)hipGetLastError();
  auto* b = new tnt_buf;
  b->ptr = ptr;
  b->len = len;
  b->dtype = dtype;
  b->device = ctx->device;
  b->owned = false;
  *out = b;
  return TNT_OK;
}

int tnt_buf_free(tnt_ctx* ctx, tnt_buf* buf) {
  if (!ctx || !buf) return TNT_E_INVALID;
  if (buf->owned) {
    if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "buffer freed during graph capture");
    // stream-ordered reuse: later allocations are used on the same stream
    pool_free(ctx, buf->ptr);
  }
  delete buf;
  return TNT_OK;
}

static int check_buf(tnt_ctx* ctx, const tnt_buf* b, const char* what) {
  if (!b) FAIL(ctx, TNT_E_INVALID, std::string("null buffer: ") + what);
  if (b->device != ctx->device) FAIL(ctx, TNT_E_DEVICE_MISMATCH, std::string(what) + " lives on another device");
  if (b->dtype != TNT_F64)
    FAIL(ctx, TNT_E_UNSUPPORTED,
         std::string(what) + " must be an FP64 payload (FP32 and complex payloads: tnt_gemm_batch only)");
  return TNT_OK;
}

// A buffer of any supported payload type on the context's device.
static int check_any_buf(tnt_ctx* ctx, const tnt_buf* b, const char* what) {
  if (!b) FAIL(ctx, TNT_E_INVALID, std::string("null buffer: ") + what);
  if (b->device != ctx->device) FAIL(ctx, TNT_E_DEVICE_MISMATCH, std::string(what) + " lives on another device");
  return TNT_OK;
}

int tnt_buf_upload(tnt_ctx* ctx, tnt_buf* buf, const void* host, size_t len) {
  if (!ctx || !buf || (!host && len)) return TNT_E_INVALID;
  if (len > buf->len) FAIL(ctx, TNT_E_SHAPE, "upload longer than buffer");
  if (int rc = check_any_buf(ctx, buf, "buf")) return rc;
  if (int rc = set_device(ctx)) return rc;
  size_t es = dtype_bytes(buf->dtype);
  HIPCHK(ctx, hipMemcpyAsync(buf->ptr, host, len * es, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TNT_OK;
}

int tnt_buf_download(tnt_ctx* ctx, const tnt_buf* buf, void* host, size_t len) {
  if (!ctx || !buf || (!host && len)) return TNT_E_INVALID;
  if (len > buf->len) FAIL(ctx, TNT_E_SHAPE, "download longer than buffer");
  if (int rc = check_any_buf(ctx, buf, "buf")) return rc;
  if (int rc = set_device(ctx)) return rc;
  size_t es = dtype_bytes(buf->dtype);
  HIPCHK(ctx, hipMemcpyAsync(host, buf->ptr, len * es, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return TNT_OK;
}

int tnt_buf_zero(tnt_ctx* ctx, tnt_buf* buf) {
  if (!ctx || !buf) return TNT_E_INVALID;
  if (int rc = check_any_buf(ctx, buf, "buf")) return rc;
  if (int rc = set_device(ctx)) return rc;
  if (buf->dtype == TNT_F32 || buf->dtype == TNT_C32)
    launch_zero_f32(static_cast<float*>(buf->ptr), (int64_t)buf->len * (buf->dtype == TNT_C32 ? 2 : 1), ctx->stream);
  else launch_fill(static_cast<double*>(buf->ptr), (int64_t)buf->len * (buf->dtype == TNT_C64 ? 2 : 1), 0.0, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  return TNT_OK;
}

int tnt_buf_fill_random(tnt_ctx* ctx, tnt_buf* buf, uint64_t seed, int normal) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, buf, "buf")) return rc;
  if (int rc = set_device(ctx)) return rc;
  if (normal) launch_fill_normal(static_cast<double*>(buf->ptr), (int64_t)buf->len, seed, ctx->stream);
  else launch_fill_unit(static_cast<double*>(buf->ptr), (int64_t)buf->len, seed, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  return TNT_OK;
}

int tnt_buf_info(const tnt_buf* buf, void** ptr, size_t* len, int* dtype, int* device) {
  if (!buf) return TNT_E_INVALID;
  if (ptr) *ptr = buf->ptr;
  if (len) *len = buf->len;
  if (dtype) *dtype = buf->dtype;
  if (device) *device = buf->device;
  return TNT_OK;
}

// Validation of tnt_gemm_batch / tnt_gemm_batch_complex: every check before
// anything is queued; fills `gj` (element offsets of the payload type).
static int gemm_batch_jobs(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                           const tnt_gemm_job* jobs, size_t n_jobs, int lhs_op, int rhs_op, int precision,
                           std::vector<GJob>& gj) {
  if (int rc = check_any_buf(ctx, dst, "dst")) return rc;
  if (int rc = check_any_buf(ctx, lhs, "lhs")) return rc;
  if (int rc = check_any_buf(ctx, rhs, "rhs")) return rc;
  if (dst->dtype != lhs->dtype || dst->dtype != rhs->dtype)
    FAIL(ctx, TNT_E_INVALID, "dst, lhs and rhs payload types differ");
  if (n_jobs && !jobs) FAIL(ctx, TNT_E_INVALID, "null job list");
  if (!valid_op(lhs_op) || !valid_op(rhs_op)) FAIL(ctx, TNT_E_INVALID, "unknown matrix op");
  if (!valid_precision(precision)) FAIL(ctx, TNT_E_INVALID, "unknown precision");
  if ((dst->dtype == TNT_F32 || dst->dtype == TNT_C32) && precision != TNT_PREC_F64 && precision != TNT_PREC_F32)
    FAIL(ctx, TNT_E_UNSUPPORTED, "split precisions need FP64 or C64 payloads");
  gj.resize(n_jobs);
  for (size_t i = 0; i < n_jobs; ++i) {
    if (int rc = make_seam_job(ctx, jobs[i], lhs_op, rhs_op, &gj[i])) return rc;
    if (!job_in_bounds(gj[i], dst->len, lhs->len, rhs->len))
      FAIL(ctx, TNT_E_OOB, "GEMM job addresses elements outside its buffers");
  }
  return TNT_OK;
}

// Complex payloads, two forms (layouts in complex.hip).  C64 with op(B) =
// identity and a real beta: R(alpha op(A)) planes in the context's workspace,
// then one real 2m x 2k x n GEMM that reads B and writes D in their real
// views (beta applied by the GEMM).  Otherwise each job becomes one real
// m x 2k x 2n job on planes of both operands -- expand pass, real block GEMM
// (alpha 1, beta 0), combine pass with the complex alpha and beta.
static int gemm_batch_complex(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                              const std::vector<GJob>& seam, int lhs_op, int rhs_op, double ar, double ai,
                              double br, double bi, int precision) {
  const bool lhs_form = dst->dtype == TNT_C64 && rhs_op == TNT_OP_IDENTITY && bi == 0.0;
  std::vector<CplxJob> cj;
  std::vector<GJob> rj;
  int64_t ws = 0;
  uint64_t a_items = 0, b_items = 0, d_items = 0;
  auto items = [](int64_t elems) { return (uint64_t)((elems + kCplxItem - 1) / kCplxItem); };
  for (const GJob& g : seam) {
    if (g.m == 0 || g.n == 0) continue;  // writes nothing
    if (2 * (int64_t)g.m > INT32_MAX || 2 * (int64_t)g.k > INT32_MAX || 2 * (int64_t)g.n > INT32_MAX)
      FAIL(ctx, TNT_E_UNSUPPORTED, "complex job: a dimension exceeds 2^30 - 1");
    const int64_t mk = (int64_t)g.m * g.k, kn = (int64_t)g.k * g.n, mn = (int64_t)g.m * g.n;
    if (lhs_form) {
      CplxJob c{};
      c.a_src = g.a_off;
      c.m = g.m;
      c.k = g.k;
      c.n = g.n;
      c.ops = ((g.ops & 1) ? CPLX_A_T : 0) | (lhs_op == TNT_OP_ADJOINT ? CPLX_A_CONJ : 0);
      c.a_ws = ws;
      ws += 4 * mk;
      c.a_item = (uint32_t)a_items;
      a_items += items(mk);
      if (a_items > UINT32_MAX) FAIL(ctx, TNT_E_UNSUPPORTED, "complex batch too large for one pass");
      cj.push_back(c);
      GJob r{};
      r.a_off = c.a_ws;          // workspace
      r.b_off = 2 * g.b_off;     // rhs payload, real view
      r.d_off = 2 * g.d_off;     // dst payload, real view
      r.m = 2 * g.m;
      r.k = 2 * g.k;
      r.n = g.n;
      r.lda = (g.ops & 1) ? std::max(r.k, 1) : std::max(r.m, 1);
      r.ldb = std::max(r.k, 1);
      r.ldd = std::max(r.m, 1);
      r.ops = g.ops & 1;
      rj.push_back(r);
      continue;
    }
    CplxJob c{};
    c.a_src = g.a_off;
    c.b_src = g.b_off;
    c.d_dst = g.d_off;
    c.m = g.m;
    c.k = g.k;
    c.n = g.n;
    c.ops = ((g.ops & 1) ? CPLX_A_T : 0) | (lhs_op == TNT_OP_ADJOINT ? CPLX_A_CONJ : 0) |
            ((g.ops & 2) ? CPLX_B_T : 0) | (rhs_op == TNT_OP_ADJOINT ? CPLX_B_CONJ : 0);
    c.a_ws = ws;
    c.b_ws = c.a_ws + 2 * mk;
    c.c_ws = c.b_ws + 4 * kn;
    ws = c.c_ws + 2 * mn;
    c.a_item = (uint32_t)a_items;
    c.b_item = (uint32_t)b_items;
    c.d_item = (uint32_t)d_items;
    a_items += items(mk);
    b_items += items(kn);
    d_items += items(mn);
    if (a_items + b_items > UINT32_MAX || d_items > UINT32_MAX)
      FAIL(ctx, TNT_E_UNSUPPORTED, "complex batch too large for one pass");
    cj.push_back(c);
    GJob r{};
    r.a_off = c.a_ws;
    r.b_off = c.b_ws;
    r.d_off = c.c_ws;
    r.m = g.m;
    r.k = 2 * g.k;
    r.n = 2 * g.n;
    r.lda = (g.ops & 1) ? std::max(r.k, 1) : std::max(r.m, 1);
    r.ldb = (g.ops & 2) ? std::max(r.n, 1) : std::max(r.k, 1);
    r.ldd = std::max(r.m, 1);
    r.ops = g.ops;
    rj.push_back(r);
  }
  if (cj.empty()) return TNT_OK;
  if (int rc = set_device(ctx)) return rc;
  begin_call(ctx);
  const size_t bytes = 8 * (size_t)ws;
  if (ctx->cplx_ws_bytes < bytes) {
    if (ctx->capturing)
      FAIL(ctx, TNT_E_INVALID, "complex GEMM workspace must grow during graph capture; run the call once before capturing");
    if (ctx->cplx_ws) retire(ctx, ctx->cplx_ws);
    ctx->cplx_ws = nullptr;
    ctx->cplx_ws_bytes = 0;
    if (int rc = pool_alloc(ctx, bytes, &ctx->cplx_ws)) return rc;
    ctx->cplx_ws_bytes = bytes;
  }
  void* dcj = nullptr;
  int rc = cached_table(ctx, cj.data(), sizeof(CplxJob) * cj.size(), "CJ", &dcj);
  GemmTable t;
  build_gemm_table(t, rj, ctx->knobs);
  void* dj = nullptr;
  void* dt = nullptr;
  if (!rc) rc = cached_table(ctx, t.host_jobs.data(), sizeof(GJob) * std::max<size_t>(t.host_jobs.size(), 1), "J", &dj);
  if (!rc) rc = cached_table(ctx, t.host_tiles.data(), sizeof(GTile) * std::max<size_t>(t.host_tiles.size(), 1), "T", &dt);
  if (rc) return rc;
  t.jobs = static_cast<GJob*>(dj);
  t.tiles = static_cast<GTile*>(dt);
  const bool c32 = dst->dtype == TNT_C32;
  double* wsp = static_cast<double*>(ctx->cplx_ws);
  const CplxJob* jobs = static_cast<const CplxJob*>(dcj);
  if (lhs_form) {
    launch_cplx_expand_lhs(jobs, (int)cj.size(), (uint32_t)a_items, lhs->ptr, wsp, ar, ai, ctx->stream);
    if (int lrc = launch_status(ctx)) return lrc;
    return run_gemm_table(ctx, t, (uintptr_t)dst->ptr, (uintptr_t)wsp, (uintptr_t)rhs->ptr, 1.0, br, precision);
  }
  launch_cplx_expand(jobs, (int)cj.size(), (uint32_t)a_items, (uint32_t)b_items, lhs->ptr, rhs->ptr, wsp, c32,
                     ctx->stream);
  if (int lrc = launch_status(ctx)) return lrc;
  if ((rc = run_gemm_table(ctx, t, (uintptr_t)wsp, (uintptr_t)wsp, (uintptr_t)wsp, 1.0, 0.0, precision))) return rc;
  launch_cplx_combine(jobs, (int)cj.size(), (uint32_t)d_items, wsp, dst->ptr, ar, ai, br, bi, c32, ctx->stream);
  return launch_status(ctx);
}

int tnt_gemm_batch(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                   const tnt_gemm_job* jobs, size_t n_jobs, const size_t* runs, size_t n_runs,
                   int lhs_op, int rhs_op, double alpha, double beta, int precision) {
  return tnt_gemm_batch_complex(ctx, dst, lhs, rhs, jobs, n_jobs, runs, n_runs, lhs_op, rhs_op, alpha, 0.0, beta,
                                0.0, precision);
}

int tnt_gemm_batch_complex(tnt_ctx* ctx, tnt_buf* dst, const tnt_buf* lhs, const tnt_buf* rhs,
                           const tnt_gemm_job* jobs, size_t n_jobs, const size_t* runs, size_t n_runs,
                           int lhs_op, int rhs_op, double alpha_re, double alpha_im, double beta_re,
                           double beta_im, int precision) {
  (void)runs;
  (void)n_runs;
  if (!ctx) return TNT_E_INVALID;
  std::vector<GJob> gj;
  if (int rc = gemm_batch_jobs(ctx, dst, lhs, rhs, jobs, n_jobs, lhs_op, rhs_op, precision, gj)) return rc;
  if (is_complex(dst->dtype))
    return gemm_batch_complex(ctx, dst, lhs, rhs, gj, lhs_op, rhs_op, alpha_re, alpha_im, beta_re, beta_im, precision);
  if (alpha_im != 0.0 || beta_im != 0.0) FAIL(ctx, TNT_E_INVALID, "complex alpha or beta on real payloads");
  const double alpha = alpha_re, beta = beta_re;
  if (int rc = set_device(ctx)) return rc;
  begin_call(ctx);
  GemmTable t;
  build_gemm_table(t, gj, ctx->knobs);
  void* dj = nullptr;
  void* dt = nullptr;
  int rc = cached_table(ctx, t.host_jobs.data(), sizeof(GJob) * std::max<size_t>(t.host_jobs.size(), 1), "J", &dj);
  if (!rc) rc = cached_table(ctx, t.host_tiles.data(), sizeof(GTile) * std::max<size_t>(t.host_tiles.size(), 1), "T", &dt);
  if (rc) return rc;
  t.jobs = static_cast<GJob*>(dj);
  t.tiles = static_cast<GTile*>(dt);
  return run_gemm_table(ctx, t, (uintptr_t)dst->ptr, (uintptr_t)lhs->ptr, (uintptr_t)rhs->ptr, alpha, beta,
                        precision, dst->dtype);
}

int tnt_gemm_grouped(tnt_ctx* ctx, const tnt_gemm_ptr_job* jobs, size_t n_jobs, double alpha,
                     double beta, int precision) {
  if (!ctx) return TNT_E_INVALID;
  if (n_jobs && !jobs) FAIL(ctx, TNT_E_INVALID, "null job list");
  if (!valid_precision(precision)) FAIL(ctx, TNT_E_INVALID, "unknown precision");
  std::vector<GJob> gj(n_jobs);
  for (size_t i = 0; i < n_jobs; ++i) {
    const tnt_gemm_ptr_job& j = jobs[i];
    if (!j.dst || !j.lhs || !j.rhs || j.rows < 0 || j.cols < 0 || j.contracted < 0)
      FAIL(ctx, TNT_E_INVALID, "invalid grouped job");
    if (!valid_op(j.lhs_op) || !valid_op(j.rhs_op)) FAIL(ctx, TNT_E_INVALID, "unknown matrix op");
    if (j.rows > INT32_MAX || j.cols > INT32_MAX || j.contracted > INT32_MAX || j.ldd > INT32_MAX ||
        j.ldl > INT32_MAX || j.ldr > INT32_MAX)
      FAIL(ctx, TNT_E_UNSUPPORTED, "grouped job dimension or leading dimension exceeds 2^31 - 1");
    if (((uintptr_t)j.dst | (uintptr_t)j.lhs | (uintptr_t)j.rhs) & 7)
      FAIL(ctx, TNT_E_INVALID, "grouped job pointers must be 8-byte aligned");
    GJob g{};
    g.d_off = (int64_t)((uintptr_t)j.dst / 8);
    g.a_off = (int64_t)((uintptr_t)j.lhs / 8);
    g.b_off = (int64_t)((uintptr_t)j.rhs / 8);
    g.m = (int32_t)j.rows;
    g.k = (int32_t)j.contracted;
    g.n = (int32_t)j.cols;
    g.lda = (int32_t)j.ldl;
    g.ldb = (int32_t)j.ldr;
    g.ldd = (int32_t)j.ldd;
    g.ops = (op_transposes(j.lhs_op) ? 1 : 0) | (op_transposes(j.rhs_op) ? 2 : 0);
    int64_t ar = (g.ops & 1) ? g.k : g.m, br = (g.ops & 2) ? g.n : g.k;
    if (g.lda < std::max<int64_t>(1, ar) || g.ldb < std::max<int64_t>(1, br) || g.ldd < std::max<int32_t>(1, g.m))
      FAIL(ctx, TNT_E_SHAPE, "leading dimension smaller than the row count");
    gj[i] = g;
  }
  if (int rc = set_device(ctx)) return rc;
  begin_call(ctx);
  GemmTable t;
  build_gemm_table(t, gj, ctx->knobs);
  void* dj = nullptr;
  void* dt = nullptr;
  int rc = cached_table(ctx, t.host_jobs.data(), sizeof(GJob) * std::max<size_t>(t.host_jobs.size(), 1), "PJ", &dj);
  if (!rc) rc = cached_table(ctx, t.host_tiles.data(), sizeof(GTile) * std::max<size_t>(t.host_tiles.size(), 1), "PT", &dt);
  if (rc) return rc;
  t.jobs = static_cast<GJob*>(dj);
  t.tiles = static_cast<GTile*>(dt);
  return run_gemm_table(ctx, t, 0, 0, 0, alpha, beta, precision);
}

int tnt_plan_upload(tnt_ctx* ctx, const tnt_transform_desc* desc, tnt_plan** out) {
  if (!ctx || !out) return TNT_E_INVALID;
  if (int rc = set_device(ctx)) return rc;
  return build_plan(ctx, desc, out);
}

int tnt_plan_destroy(tnt_ctx* ctx, tnt_plan* plan) {
  if (!ctx || !plan) return TNT_E_INVALID;
  (void)hipStreamSynchronize(ctx->stream);
  for (auto a : plan->allocs) pool_free(ctx, a);
  delete plan;
  return TNT_OK;
}

int tnt_tree_transform(tnt_ctx* ctx, const tnt_plan* plan, tnt_buf* dst, const tnt_buf* src,
                       double alpha, double beta, int overwrite) {
  if (!ctx || !plan) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, dst, "dst")) return rc;
  if (int rc = check_buf(ctx, src, "src")) return rc;
  if (plan->device != ctx->device) FAIL(ctx, TNT_E_DEVICE_MISMATCH, "plan uploaded to another device");
  if ((int64_t)dst->len < plan->dst_len) FAIL(ctx, TNT_E_SHAPE, "destination shorter than the plan's destination payload");
  if ((int64_t)src->len < plan->src_len) FAIL(ctx, TNT_E_SHAPE, "source shorter than the plan's source payload");
  if (dst->ptr == src->ptr && plan->dst_len > 0) FAIL(ctx, TNT_E_INVALID, "in-place tree transform is not supported");
  if (int rc = set_device(ctx)) return rc;
  run_plan(ctx, plan, static_cast<double*>(dst->ptr), static_cast<const double*>(src->ptr), alpha, beta, overwrite);
  return launch_status(ctx);
}

// Element maps of the tiny route.  Layouts were bounds-checked by build_plan.
// Visits every element of a layout pair in column-major order of the shape.
extern "C++" template <class F>
static void for_each_element(const tnt_layout& ld, const tnt_layout& ls, F&& f) {
  int64_t n = 1;
  for (int i = 0; i < ld.rank; ++i) n *= ld.shape[i];
  if (n <= 0) return;
  int64_t c[TNT_MAX_RANK] = {0};
  for (int64_t e = 0; e < n; ++e) {
    int64_t od = ld.offset, os = ls.offset;
    for (int i = 0; i < ld.rank; ++i) {
      od += c[i] * ld.strides[i];
      os += c[i] * ls.strides[i];
    }
    f(od, os);
    for (int i = 0; i < ld.rank; ++i) {
      if (++c[i] < ld.shape[i]) break;
      c[i] = 0;
    }
  }
}

static bool single_only(const tnt_transform_desc* d) {
  for (size_t i = 0; i < d->n_blocks; ++i)
    if (d->blocks[i].kind != TNT_BLOCK_SINGLE) return false;
  return true;
}

// Gather map of an operand transform replayed with overwrite: transformed
// element e = c * src[idx] (c = 0, idx = 0 where no block writes e).
static bool tiny_gather_map(const tnt_transform_desc* d, std::vector<TinyEntry>& m) {
  if (d->src_len < 1) return false;
  m.assign((size_t)d->dst_len, TinyEntry{0, 0.0});
  for (size_t i = 0; i < d->n_blocks; ++i) {
    const tnt_transform_block& b = d->blocks[i];
    const double c = d->coeffs[b.coeff];
    for_each_element(d->layouts[b.dst_layout], d->layouts[b.src_layout],
                     [&](int64_t od, int64_t os) { m[(size_t)od] = TinyEntry{os, c}; });
  }
  return true;
}

// Scatter map of the output transform: core element e goes to dst[idx] scaled
// by c (idx = -1: read by no block).  Fails when a core element feeds two
// destination elements.  `fix` receives the destination elements of inactive
// layouts (beta * dst).
static bool tiny_scatter_map(const tnt_transform_desc* d, std::vector<TinyEntry>& m, std::vector<int64_t>& fix) {
  m.assign((size_t)d->src_len, TinyEntry{-1, 0.0});
  bool ok = true;
  for (size_t i = 0; i < d->n_blocks && ok; ++i) {
    const tnt_transform_block& b = d->blocks[i];
    const double c = d->coeffs[b.coeff];
    for_each_element(d->layouts[b.dst_layout], d->layouts[b.src_layout], [&](int64_t od, int64_t os) {
      if (m[(size_t)os].idx >= 0) ok = false;
      m[(size_t)os] = TinyEntry{od, c};
    });
  }
  for (size_t i = 0; i < d->n_inactive && ok; ++i) {
    const tnt_layout& l = d->layouts[d->inactive[i]];
    for_each_element(l, l, [&](int64_t od, int64_t) { fix.push_back(od); });
  }
  return ok;
}

// Decides and builds the tiny route of an uploaded artifact (after the general
// route, which stays for the other precisions).  Returns non-zero only on an
// allocation or copy failure.
// The output transform folded into the tile GEMM's epilogue: eligible when
// it is Single-only with no inactive layouts, the core has no uncovered
// region, and every block maps whole core columns -- its core-side rows are
// all rows of its job, contiguous (first row axis stride 1, each next axis
// the product of the previous extents), and its destination-side row
// strides follow the same nesting -- so that core element (row, col) goes to
// dst[off(col) + row * rs(col)] scaled by the block coefficient, and the
// blocks together write every destination element once.  Not eligible:
// nothing is built (the contraction keeps transform-after-GEMM).
// TENET_AMD_OUT_EPILOGUE=0 at upload keeps every contraction on the
// separate output pass (A/B runs and tests).
static bool out_epilogue_enabled() {
  const char* e = std::getenv("TENET_AMD_OUT_EPILOGUE");
  return !(e && e[0] == '0');
}

// The fused narrow route gathers its coefficients inside the recoupling
// launch where it can (TENET_AMD_GATHER_FOLD=0: a gather launch first, for
// A/B runs).
static bool gather_fold_enabled() {
  const char* e = std::getenv("TENET_AMD_GATHER_FOLD");
  return !(e && e[0] == '0');
}

// Operand transforms of a contraction run on two streams when each moves at
// least this many bytes (TENET_AMD_FORK_TRANSFORMS=0: always one stream, for
// A/B runs).
constexpr double kForkPlanBytes = 64.0 * (1 << 20);
static bool fork_enabled() {
  const char* e = std::getenv("TENET_AMD_FORK_TRANSFORMS");
  return !(e && e[0] == '0');
}

// Uncovered core regions as GEMM jobs with nothing contracted unless
// TENET_AMD_REGION_JOBS=0 (then a region pass of their own; for A/B runs).
static bool region_jobs_enabled() {
  const char* e = std::getenv("TENET_AMD_REGION_JOBS");
  return !(e && e[0] == '0');
}

static int build_colmap(tnt_ctx* ctx, const tnt_contract_desc* d, const std::vector<GJob>& gj, tnt_artifact* a) {
  const tnt_transform_desc* t = d->out;
  if (!t || !out_epilogue_enabled() || t->n_inactive || d->n_uncovered) return TNT_OK;
  std::vector<int64_t> start(gj.size() + 1, 0);
  for (size_t j = 0; j < gj.size(); ++j) start[j + 1] = start[j] + std::max(gj[j].n, 0);
  std::vector<ColMapEntry> map((size_t)start.back(), ColMapEntry{-1, 0, 0.0});
  std::vector<std::pair<int64_t, size_t>> starts;  // core offset -> job
  for (size_t j = 0; j < gj.size(); ++j)
    if (gj[j].m > 0 && gj[j].n > 0) starts.push_back({gj[j].d_off, j});
  std::sort(starts.begin(), starts.end());
  int64_t written = 0;
  for (size_t bi = 0; bi < t->n_blocks; ++bi) {
    const tnt_transform_block& b = t->blocks[bi];
    if (b.kind != TNT_BLOCK_SINGLE || b.dst_count != 1 || b.src_count != 1 || b.coeff >= t->n_coeffs ||
        b.dst_layout >= t->n_layouts || b.src_layout >= t->n_layouts)
      return TNT_OK;
    const tnt_layout& Od = t->layouts[b.dst_layout];
    const tnt_layout& Oc = t->layouts[b.src_layout];
    if (Od.rank != Oc.rank || Oc.rank < 0 || Oc.rank > TNT_MAX_RANK) return TNT_OK;
    int64_t cnt = 1, dhi = Od.offset;
    for (int i = 0; i < Oc.rank; ++i) {
      if (Od.shape[i] != Oc.shape[i] || Oc.shape[i] < 0 || Od.strides[i] < 0) return TNT_OK;
      cnt *= Oc.shape[i];
      if (Oc.shape[i] > 0) dhi += (Oc.shape[i] - 1) * Od.strides[i];
    }
    if (cnt == 0) continue;
    if (Od.offset < 0 || dhi >= t->dst_len) return TNT_OK;
    auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(Oc.offset, SIZE_MAX));
    if (it == starts.begin()) return TNT_OK;
    --it;
    const GJob& J = gj[it->second];
    const int64_t rel = Oc.offset - J.d_off, m = J.m;
    if (rel < 0 || rel >= m * (int64_t)J.n || rel % m) return TNT_OK;  // must start at row 0
    const int64_t c0 = rel / m;
    std::vector<int> rows, cols;
    for (int i = 0; i < Oc.rank; ++i) {
      if (Oc.shape[i] == 1) continue;
      if (Oc.strides[i] < m) rows.push_back(i);
      else if (Oc.strides[i] % m) return TNT_OK;
      else cols.push_back(i);
    }
    std::sort(rows.begin(), rows.end(), [&](int x, int y) { return Oc.strides[x] < Oc.strides[y]; });
    int64_t prod = 1;
    const int64_t rs = rows.empty() ? 0 : Od.strides[rows[0]];
    for (int i : rows) {
      if (Oc.strides[i] != prod || Od.strides[i] != rs * prod) return TNT_OK;
      prod *= Oc.shape[i];
    }
    if (prod != m) return TNT_OK;  // the block covers every row of its job
    const double c = t->coeffs[b.coeff];
    std::vector<int64_t> idx(cols.size(), 0);
    while (true) {
      int64_t col = c0, off = Od.offset;
      for (size_t u = 0; u < cols.size(); ++u) {
        col += idx[u] * (Oc.strides[cols[u]] / m);
        off += idx[u] * Od.strides[cols[u]];
      }
      if (col >= J.n) return TNT_OK;
      ColMapEntry& e = map[(size_t)(start[it->second] + col)];
      if (e.off >= 0) return TNT_OK;  // a core column mapped twice
      e = ColMapEntry{off, rs, c};
      size_t u = 0;
      for (; u < cols.size(); ++u) {
        if (++idx[u] < Oc.shape[cols[u]]) break;
        idx[u] = 0;
      }
      if (u == cols.size()) break;
    }
    written += cnt;
  }
  if (written != t->dst_len) return TNT_OK;  // some destination element would be left unwritten
  start.pop_back();
  if (int rc = upload_vec(ctx, start, &a->cmap_start, a->allocs)) return rc;
  return upload_vec(ctx, map, &a->cmap, a->allocs);
}

static int build_tiny_route(tnt_ctx* ctx, const tnt_contract_desc* d, const std::vector<GJob>& gj,
                            tnt_artifact* a) {
  const double max_flops = tiny_max_flops();
  if (max_flops <= 0) return TNT_OK;
  double flops = 0;
  for (const auto& j : gj) flops += 2.0 * j.m * (double)j.k * j.n;
  if (flops > max_flops) return TNT_OK;
  const int64_t dst_len = d->out ? d->out->dst_len : d->core_len;
  for (int64_t len : {d->lhs_len, d->rhs_len, d->core_len, dst_len,
                      d->lhs ? d->lhs->src_len : int64_t(0), d->rhs ? d->rhs->src_len : int64_t(0)})
    if (len > kTinyMaxElems) return TNT_OK;
  if ((d->lhs && !single_only(d->lhs)) || (d->rhs && !single_only(d->rhs)) || (d->out && !single_only(d->out)))
    return TNT_OK;
  std::vector<TinyEntry> lm, rm, om;
  std::vector<int64_t> fix;
  if (d->lhs && !tiny_gather_map(d->lhs, lm)) return TNT_OK;
  if (d->rhs && !tiny_gather_map(d->rhs, rm)) return TNT_OK;
  if (d->out) {
    if (!tiny_scatter_map(d->out, om, fix)) return TNT_OK;
    // core elements of uncovered regions are zero: their destinations get beta * dst
    for (size_t i = 0; i < d->n_uncovered; ++i)
      for (int64_t e = d->uncovered[i].offset; e < d->uncovered[i].offset + d->uncovered[i].len; ++e)
        if (om[(size_t)e].idx >= 0) fix.push_back(om[(size_t)e].idx);
  } else {
    for (size_t i = 0; i < d->n_uncovered; ++i)
      for (int64_t e = d->uncovered[i].offset; e < d->uncovered[i].offset + d->uncovered[i].len; ++e)
        fix.push_back(e);
  }
  std::vector<GTile> tiles;
  for (size_t ji = 0; ji < gj.size(); ++ji) {
    const GJob& j = gj[ji];
    for (int tm = 0; tm < (j.m + 31) / 32; ++tm)
      for (int tn = 0; tn < (j.n + 31) / 32; ++tn) tiles.push_back(GTile{(int32_t)ji, tm, tn, TILE_SMALL});
  }
  int rc = TNT_OK;
  if (!rc && d->lhs) rc = upload_vec(ctx, lm, &a->t_lmap, a->allocs);
  if (!rc && d->rhs) rc = upload_vec(ctx, rm, &a->t_rmap, a->allocs);
  if (!rc && d->out) rc = upload_vec(ctx, om, &a->t_omap, a->allocs);
  if (!rc) rc = upload_vec(ctx, fix, &a->t_fix, a->allocs);
  if (!rc) rc = upload_vec(ctx, gj, &a->t_jobs, a->allocs);
  if (!rc) rc = upload_vec(ctx, tiles, &a->t_tiles, a->allocs);
  if (rc) return rc;
  a->t_n_fix = (uint32_t)fix.size();
  a->t_n_tiles = (uint32_t)tiles.size();
  a->t_writes = fix.size();
  for (const GJob& j : gj)
    for (int col = 0; col < j.n; ++col)
      for (int row = 0; row < j.m; ++row)
        if (om.empty() || om[(size_t)(j.d_off + row + (int64_t)col * j.ldd)].idx >= 0) ++a->t_writes;
  a->t_flops = flops;
  a->tiny = true;
  return TNT_OK;
}

int tnt_contract_upload(tnt_ctx* ctx, const tnt_contract_desc* d, tnt_artifact** out) {
  if (!ctx || !d || !out) return TNT_E_INVALID;
  if (!valid_op(d->lhs_op) || !valid_op(d->rhs_op)) FAIL(ctx, TNT_E_INVALID, "unknown matrix op");
  if (d->n_jobs && !d->jobs) FAIL(ctx, TNT_E_INVALID, "null job list");
  if (d->n_uncovered && !d->uncovered) FAIL(ctx, TNT_E_INVALID, "null uncovered-region list");
  if (d->lhs_len < 0 || d->rhs_len < 0 || d->core_len < 0) FAIL(ctx, TNT_E_SHAPE, "negative operand length");
  std::vector<GJob> gj(d->n_jobs);
  for (size_t i = 0; i < d->n_jobs; ++i) {
    if (int rc = make_seam_job(ctx, d->jobs[i], d->lhs_op, d->rhs_op, &gj[i])) return rc;
    if (!job_in_bounds(gj[i], (size_t)d->core_len, (size_t)d->lhs_len, (size_t)d->rhs_len))
      FAIL(ctx, TNT_E_OOB, "core GEMM job addresses elements outside its operands");
  }
  if (int rc = set_device(ctx)) return rc;
  auto* a = new tnt_artifact;
  a->device = ctx->device;
  a->lhs_len = d->lhs_len;
  a->rhs_len = d->rhs_len;
  a->core_len = d->core_len;
  a->lhs_op = d->lhs_op;
  a->rhs_op = d->rhs_op;
  auto fail = [&](int rc) {
    tnt_contract_destroy(ctx, a);
    return rc;
  };
  if (d->lhs) {
    if (int rc = build_plan(ctx, d->lhs, &a->lhs)) return fail(rc);
    if (a->lhs->dst_len != d->lhs_len) return fail((ctx->err = "lhs transform length mismatch", TNT_E_SHAPE));
  }
  if (d->rhs) {
    if (int rc = build_plan(ctx, d->rhs, &a->rhs)) return fail(rc);
    if (a->rhs->dst_len != d->rhs_len) return fail((ctx->err = "rhs transform length mismatch", TNT_E_SHAPE));
  }
  if (d->out) {
    if (int rc = build_plan(ctx, d->out, &a->out)) return fail(rc);
    if (a->out->src_len != d->core_len) return fail((ctx->err = "output transform length mismatch", TNT_E_SHAPE));
  }
  if (int rc = build_colmap(ctx, d, gj, a)) return fail(rc);
  // Narrow outputs (n <= 16) go to workgroup tiles, in the launch of the
  // other tiles, when they are at most 2 % of the flops: their padded tiles
  // cost less than a launch of the narrow-output kernel of their own (and an
  // output transform folded into the epilogue needs workgroup tiles).
  // TENET_AMD_NARROW_ON_TILES=0 keeps them on that kernel unless the output
  // transform is folded (for A/B runs).
  bool no_skinny = false;
  const char* nt_env = std::getenv("TENET_AMD_NARROW_ON_TILES");
  if (a->cmap || !(nt_env && nt_env[0] == '0')) {
    double all = 0.0, narrow = 0.0;
    for (const GJob& j : gj) {
      const double f = 2.0 * double(std::max(j.m, 0)) * double(std::max(j.n, 0)) * double(std::max(j.k, 0));
      all += f;
      if (j.n <= 16) narrow += f;
    }
    no_skinny = narrow <= 0.02 * all;
  }
  for (size_t i = 0; i < d->n_uncovered; ++i)
    if (d->uncovered[i].offset < 0 || d->uncovered[i].len < 0 ||
        d->uncovered[i].offset + d->uncovered[i].len > d->core_len) {
      ctx->err = "uncovered region out of range";
      return fail(TNT_E_OOB);
    }
  // Without an output transform the uncovered regions (dst = beta dst) are
  // jobs of the GEMM table with nothing contracted, in the launch of the
  // contraction's own tiles (they are a few hundred elements, and a launch of
  // their own costs more than their bytes); with one they stay regions of
  // the core scratch, zeroed before the GEMM.
  std::vector<GJob> table_jobs = gj;
  std::vector<int64_t> uo, ul;
  if (!d->out && region_jobs_enabled()) {
    region_jobs(d->uncovered, d->n_uncovered, gj.empty() ? 0 : gj[0].ops, table_jobs);
  } else {
    for (size_t i = 0; i < d->n_uncovered; ++i) {
      uo.push_back(d->uncovered[i].offset);
      ul.push_back(d->uncovered[i].len);
    }
  }
  build_gemm_table(a->gemm, table_jobs, ctx->knobs, -1, no_skinny);
  if (int rc = upload_gemm_table(ctx, a->gemm, a->allocs)) return fail(rc);
  a->n_unc = (uint32_t)uo.size();
  for (const GJob& j : table_jobs) a->core_writes += (uint64_t)std::max(j.m, 0) * (uint64_t)std::max(j.n, 0);
  for (int64_t l : ul) a->core_writes += (uint64_t)l;
  if (int rc = upload_vec(ctx, uo, &a->unc_off, a->allocs)) return fail(rc);
  if (int rc = upload_vec(ctx, ul, &a->unc_len, a->allocs)) return fail(rc);
  if (fusion_enabled()) {
    tnt_fuse::NarrowFusion nf;
    if (tnt_fuse::plan_narrow_fusion(d, nf)) {
      if (build_plan(ctx, &nf.desc, &a->fused) != TNT_OK) {
        a->fused = nullptr;  // not expressible as a plan: keep the general route
        ctx->err.clear();
      } else {
        if (nf.gather_ptr.size() - 1 > UINT32_MAX) return fail((ctx->err = "fused plan has too many coefficients", TNT_E_UNSUPPORTED));
        a->n_gather = (uint32_t)(nf.gather_ptr.size() - 1);
        if (int rc = upload_vec(ctx, nf.gather_ptr, &a->g_ptr, a->allocs)) return fail(rc);
        if (int rc = upload_vec(ctx, nf.gather_idx, &a->g_idx, a->allocs)) return fail(rc);
        if (int rc = upload_vec(ctx, nf.gather_scale, &a->g_scale, a->allocs)) return fail(rc);
        const auto& cc = a->fused->d;
        bool fold = cc.cls_count[XC_MULTI4] > 0;
        for (int c = 0; c < XC_COUNT; ++c)
          if (c != XC_MULTI4 && cc.cls_count[c]) fold = false;
        for (size_t bi = 0; fold && bi < nf.desc.n_blocks; ++bi)
          if ((size_t)nf.desc.blocks[bi].dst_count * nf.desc.blocks[bi].src_count > (size_t)kMaxGatherCoef) fold = false;
        a->gather_fold = fold;
        *out = a;  // the fused route needs no scratch payloads
        return TNT_OK;
      }
    }
  }
  if (int rc = build_tiny_route(ctx, d, gj, a)) return fail(rc);
  void* p = nullptr;
  if (a->lhs) {
    if (int rc = pool_alloc(ctx, 8 * (size_t)std::max<int64_t>(a->lhs_len, 1), &p)) return fail(rc);
    a->allocs.push_back(p);
    a->scratch_l = static_cast<double*>(p);
  }
  if (a->rhs) {
    if (int rc = pool_alloc(ctx, 8 * (size_t)std::max<int64_t>(a->rhs_len, 1), &p)) return fail(rc);
    a->allocs.push_back(p);
    a->scratch_r = static_cast<double*>(p);
  }
  if (a->out) {
    if (int rc = pool_alloc(ctx, 8 * (size_t)std::max<int64_t>(a->core_len, 1), &p)) return fail(rc);
    a->allocs.push_back(p);
    a->scratch_c = static_cast<double*>(p);
  }
  *out = a;
  return TNT_OK;
}

int tnt_contract_set_tile_config(tnt_ctx* ctx, tnt_artifact* a, int cfg) {
  if (!ctx || !a) return TNT_E_INVALID;
  if (cfg < -1 || cfg >= TILE_CFG_COUNT) FAIL(ctx, TNT_E_INVALID, "unknown tile configuration");
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "tile configuration cannot change during graph capture");
  if (int rc = set_device(ctx)) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<GJob> jobs = a->gemm.host_jobs;
  GemmTable t;
  build_gemm_table(t, jobs, ctx->knobs, cfg, a->gemm.no_skinny);
  std::vector<void*> allocs;
  if (int rc = upload_gemm_table(ctx, t, allocs)) {
    for (auto p : allocs) pool_free(ctx, p);
    return rc;
  }
  // release the previous tables (a live graph may have recorded them)
  auto drop = [&](void* p) {
    auto it = std::find(a->allocs.begin(), a->allocs.end(), p);
    if (it != a->allocs.end()) {
      retire(ctx, p);
      a->allocs.erase(it);
    }
  };
  drop(a->gemm.jobs);
  drop(a->gemm.tiles);
  a->gemm = t;
  a->allocs.insert(a->allocs.end(), allocs.begin(), allocs.end());
  return TNT_OK;
}

int tnt_ctx_set_shadow(tnt_ctx* ctx, int on) {
  if (!ctx) return TNT_E_INVALID;
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "the shadow check cannot change during graph capture");
  if (int rc = set_device(ctx)) return rc;
  if (on) {
    if (!ctx->shadow_stats) HIPCHK(ctx, hipMalloc(&ctx->shadow_stats, 2 * sizeof(unsigned long long)));
    launch_fill(reinterpret_cast<double*>(ctx->shadow_stats), 2, 0.0, ctx->stream);  // all-zero bits
    HIPCHK(ctx, hipGetLastError());
    ctx->shadow_expected = ctx->shadow_passes = 0;
  }
  ctx->shadow = on != 0;
  return TNT_OK;
}

int tnt_ctx_shadow_stats(tnt_ctx* ctx, uint64_t* writes, uint64_t* expected, uint64_t* doubles, uint64_t* passes) {
  if (!ctx || !writes || !expected || !doubles || !passes) return TNT_E_INVALID;
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "tnt_ctx_shadow_stats synchronises and cannot be captured");
  if (int rc = set_device(ctx)) return rc;
  unsigned long long st[2] = {0ull, 0ull};
  if (ctx->shadow_stats) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(st, ctx->shadow_stats, sizeof(st), hipMemcpyDeviceToHost));
  }
  *writes = st[0];
  *doubles = st[1];
  *expected = ctx->shadow_expected;
  *passes = ctx->shadow_passes;
  return TNT_OK;
}

int tnt_contract_route(const tnt_artifact* a, int* route) {
  if (!a || !route) return TNT_E_INVALID;
  *route = a->fused ? TNT_ROUTE_FUSED_NARROW : a->tiny ? TNT_ROUTE_TINY_GATHERED : TNT_ROUTE_TRANSFORM_GEMM;
  return TNT_OK;
}

int tnt_contract_destroy(tnt_ctx* ctx, tnt_artifact* a) {
  if (!ctx || !a) return TNT_E_INVALID;
  (void)hipStreamSynchronize(ctx->stream);
  if (a->lhs) tnt_plan_destroy(ctx, a->lhs);
  if (a->rhs) tnt_plan_destroy(ctx, a->rhs);
  if (a->out) tnt_plan_destroy(ctx, a->out);
  if (a->fused) tnt_plan_destroy(ctx, a->fused);
  for (auto p : a->allocs) pool_free(ctx, p);
  delete a;
  return TNT_OK;
}

int tnt_contract(tnt_ctx* ctx, const tnt_artifact* a, tnt_buf* dst, const tnt_buf* lhs,
                 const tnt_buf* rhs, double alpha, double beta, int precision) {
  if (!ctx || !a) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, dst, "dst")) return rc;
  if (int rc = check_buf(ctx, lhs, "lhs")) return rc;
  if (int rc = check_buf(ctx, rhs, "rhs")) return rc;
  if (a->device != ctx->device) FAIL(ctx, TNT_E_DEVICE_MISMATCH, "artifact uploaded to another device");
  if (!valid_precision(precision)) FAIL(ctx, TNT_E_INVALID, "unknown precision");
  int64_t need_l = a->lhs ? a->lhs->src_len : a->lhs_len;
  int64_t need_r = a->rhs ? a->rhs->src_len : a->rhs_len;
  int64_t need_d = a->out ? a->out->dst_len : a->core_len;
  if ((int64_t)lhs->len < need_l || (int64_t)rhs->len < need_r || (int64_t)dst->len < need_d)
    FAIL(ctx, TNT_E_SHAPE, "operand payload shorter than the artifact expects");
  if (dst->ptr == lhs->ptr || dst->ptr == rhs->ptr) FAIL(ctx, TNT_E_INVALID, "destination aliases an operand");
  if (int rc = set_device(ctx)) return rc;
  const double* L = static_cast<const double*>(lhs->ptr);
  const double* R = static_cast<const double*>(rhs->ptr);
  if (a->fused) {
    if (a->gather_fold && !ctx->shadow && gather_fold_enabled()) {
      // one launch: coefficients gathered per item, then the recoupling
      const tnt_plan* p = a->fused;
      XformArgs x{p->d.blocks, p->d.layouts, p->d.items, p->d.flat, p->d.coeffs, static_cast<double*>(dst->ptr),
                  L, alpha, beta, 0};
      x.gptr = a->g_ptr;
      x.gidx = a->g_idx;
      x.gscale = a->g_scale;
      x.gsrc = R;
      TimedScope ts(ctx, 0, p->bytes_moved);
      launch_xform_multi_gather(x, p->d.cls_begin[XC_MULTI4], p->d.cls_count[XC_MULTI4], ctx->stream);
    } else {
      launch_coeff_gather(a->fused->d.coeffs, R, a->g_ptr, a->g_idx, a->g_scale, a->n_gather, ctx->stream);
      run_plan(ctx, a->fused, static_cast<double*>(dst->ptr), L, alpha, beta, 0);
    }
    return launch_status(ctx);
  }
  if (a->tiny && precision == TNT_PREC_F64) {
    TimedScope ts(ctx, 1 + precision, ctx->timing ? a->t_flops : 0.0);
    TimedScope tk(ctx, timing_kind_kernel(precision, TILE_SMALL), ctx->timing ? a->t_flops : 0.0);
    GemmArgs g{a->t_jobs, a->t_tiles, (uintptr_t)dst->ptr, (uintptr_t)L, (uintptr_t)R, alpha, beta};
    TinyArgs t{a->t_lmap, a->t_rmap, a->t_omap, a->t_fix, a->t_n_fix, 0u, static_cast<double*>(dst->ptr)};
    launch_gemm_tiny(g, t, a->t_n_tiles, ctx->stream);
    if (int rc = shadow_pass(ctx, a->out ? a->out->dst_len : a->core_len, a->t_writes, [&](uint32_t* cnt) {
          launch_shadow_tiny(a->t_jobs, a->t_tiles, a->t_n_tiles, a->t_omap, a->t_fix, a->t_n_fix, cnt, ctx->stream);
        }))
      return rc;
    return launch_status(ctx);
  }
  // Operand transforms on two streams.  Both large: the rhs one runs on the
  // side stream while the lhs one runs here (two HBM-bound passes of short
  // runs side by side move more bytes per second than one after the other:
  // 3.61 -> 3.32 ms for the SU(2) F-move's pair).  One large transform and a
  // split-INT8 table (no output transform): that transform runs on the side
  // stream and the other operand's exponent and digit passes run here beside
  // it.  The split-INT8 table joins the side stream before its GEMM itself
  // (run_gemm_table); every other path joins before the GEMM here.  Serial
  // under graph capture, per-kernel timing and the shadow check.
  const bool big_l = a->lhs && a->lhs->bytes_moved >= kForkPlanBytes;
  const bool big_r = a->rhs && a->rhs->bytes_moved >= kForkPlanBytes;
  const bool split_table = precision == TNT_PREC_SPLIT_I8 && !a->out;
  int fork_op = -1;  // operand (0 lhs, 1 rhs) whose transform runs on the side stream
  if (!ctx->capturing && !ctx->timing && !ctx->shadow && fork_enabled()) {
    if (big_l && big_r) fork_op = 1;
    else if (split_table && big_r && !a->lhs) fork_op = 1;
    else if (split_table && big_l && !a->rhs) fork_op = 0;
  }
  if (fork_op >= 0) {
    if (!ctx->side) {  // the stream and both events, or none of them
      hipStream_t st = nullptr;
      hipEvent_t ef = nullptr, ej = nullptr;
      const hipError_t e1 = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
      const hipError_t e2 = e1 == hipSuccess ? hipEventCreateWithFlags(&ef, hipEventDisableTiming) : e1;
      const hipError_t e3 = e2 == hipSuccess ? hipEventCreateWithFlags(&ej, hipEventDisableTiming) : e2;
      if (e3 != hipSuccess) {
        if (ef) (void)hipEventDestroy(ef);
        if (st) (void)hipStreamDestroy(st);
        FAIL(ctx, TNT_E_HIP, std::string("side stream for the operand transforms: ") + hipGetErrorString(e3));
      }
      ctx->side = st;
      ctx->fork_ev = ef;
      ctx->join_ev = ej;
    }
    HIPCHK(ctx, hipEventRecord(ctx->fork_ev, ctx->stream));
    HIPCHK(ctx, hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0));
    std::swap(ctx->stream, ctx->side);
    if (fork_op == 1) run_plan(ctx, a->rhs, a->scratch_r, R, 1.0, 0.0, 1);
    else run_plan(ctx, a->lhs, a->scratch_l, L, 1.0, 0.0, 1);
    std::swap(ctx->stream, ctx->side);
    HIPCHK(ctx, hipEventRecord(ctx->join_ev, ctx->side));
  }
  if (a->lhs) {
    if (fork_op != 0) run_plan(ctx, a->lhs, a->scratch_l, L, 1.0, 0.0, 1);
    L = a->scratch_l;
  }
  if (a->rhs) {
    if (fork_op != 1) run_plan(ctx, a->rhs, a->scratch_r, R, 1.0, 0.0, 1);
    R = a->scratch_r;
  }
  const bool join_in_table = fork_op >= 0 && split_table;
  if (fork_op >= 0 && !join_in_table) HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, ctx->join_ev, 0));
  int rc;
  // shadow check of the core destination: the GEMM table's tiles (FP64 / FP32
  // tables; the split precisions run their own tiles) and the uncovered regions
  auto shadow_core = [&]() -> int {
    if (precision != TNT_PREC_F64 && precision != TNT_PREC_F32) return TNT_OK;
    const uint32_t n_tiles = (uint32_t)a->gemm.host_tiles.size();
    return shadow_pass(ctx, a->core_len, a->core_writes, [&](uint32_t* cnt) {
      launch_shadow_tiles(a->gemm.jobs, a->gemm.tiles, n_tiles, cnt, ctx->stream);
      launch_shadow_regions(a->unc_off, a->unc_len, a->n_unc, cnt, ctx->stream);
    });
  };
  // the output transform folded into the epilogue: FP64 / FP32 tables on
  // workgroup tiles only (the narrow-output and one-wave kernels write the
  // core layout), not under the shadow check (it counts core writes)
  if (a->out && a->cmap && (precision == TNT_PREC_F64 || precision == TNT_PREC_F32) && !ctx->shadow &&
      !a->gemm.n_tiles_cfg[TILE_SKINNY] && !a->gemm.n_tiles_cfg[TILE_SMALL]) {
    rc = run_gemm_table(ctx, a->gemm, (uintptr_t)dst->ptr, (uintptr_t)L, (uintptr_t)R, alpha, beta, precision, TNT_F64,
                        a->cmap_start, a->cmap);
    if (rc) return rc;
    return launch_status(ctx);
  }
  if (a->out) {
    launch_scale_regions(a->scratch_c, a->unc_off, a->unc_len, a->n_unc, 0.0, ctx->stream);
    rc = run_gemm_table(ctx, a->gemm, (uintptr_t)a->scratch_c, (uintptr_t)L, (uintptr_t)R, 1.0, 0.0, precision);
    if (rc) return rc;
    if ((rc = shadow_core())) return rc;
    run_plan(ctx, a->out, static_cast<double*>(dst->ptr), a->scratch_c, alpha, beta, 0);
  } else {
    launch_scale_regions(static_cast<double*>(dst->ptr), a->unc_off, a->unc_len, a->n_unc, beta, ctx->stream);
    rc = run_gemm_table(ctx, a->gemm, (uintptr_t)dst->ptr, (uintptr_t)L, (uintptr_t)R, alpha, beta, precision,
                        TNT_F64, nullptr, nullptr, join_in_table ? ctx->side : nullptr,
                        join_in_table ? ctx->join_ev : nullptr, join_in_table ? fork_op : 1);
    if (rc) return rc;
    if ((rc = shadow_core())) return rc;
  }
  return launch_status(ctx);
}

int tnt_weighted_dot(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                     const double* weights, size_t n_regions, double* out) {
  if (!ctx || !out) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, a, "a")) return rc;
  if (int rc = check_buf(ctx, b, "b")) return rc;
  if (n_regions && (!regions || !weights)) FAIL(ctx, TNT_E_INVALID, "null region list");
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "tnt_weighted_dot returns a host value and cannot be captured");
  std::vector<int64_t> cb, ce;
  std::vector<int32_t> cr;
  std::vector<double> w(weights, weights + n_regions);
  const int64_t chunk = 1 << 16;
  for (size_t r = 0; r < n_regions; ++r) {
    const tnt_region& g = regions[r];
    if (g.offset < 0 || g.len < 0 || (size_t)(g.offset + g.len) > a->len || (size_t)(g.offset + g.len) > b->len)
      FAIL(ctx, TNT_E_OOB, "region outside the payloads");
    for (int64_t s = 0; s < g.len; s += chunk) {
      cb.push_back(g.offset + s);
      ce.push_back(g.offset + std::min(g.len, s + chunk));
      cr.push_back((int32_t)r);
    }
  }
  if (int rc = set_device(ctx)) return rc;
  if (cb.empty()) {
    *out = 0.0;
    return TNT_OK;
  }
  begin_call(ctx);
  // one packed table: begins | ends | partials, regions and weights separately
  void *dcb, *dce, *dcr, *dw;
  int rc = cached_table(ctx, cb.data(), 8 * cb.size(), "DB", &dcb);
  if (!rc) rc = cached_table(ctx, ce.data(), 8 * ce.size(), "DE", &dce);
  if (!rc) rc = cached_table(ctx, cr.data(), 4 * cr.size(), "DR", &dcr);
  if (!rc) rc = cached_table(ctx, w.data(), 8 * w.size(), "DW", &dw);
  if (rc) return rc;
  void* parts = nullptr;
  if ((rc = pool_alloc(ctx, 8 * cb.size(), &parts))) return rc;
  launch_weighted_dot_partials(static_cast<const double*>(a->ptr), static_cast<const double*>(b->ptr),
                               static_cast<int64_t*>(dcb), static_cast<int64_t*>(dce),
                               static_cast<int32_t*>(dcr), static_cast<double*>(dw),
                               static_cast<double*>(parts), (uint32_t)cb.size(), ctx->stream);
  launch_sum_partials(static_cast<double*>(parts), (uint32_t)cb.size(), ctx->dev_scalar, ctx->stream);
  HIPCHK(ctx, hipMemcpyAsync(ctx->host_scalar, ctx->dev_scalar, 8, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  pool_free(ctx, parts);
  *out = *ctx->host_scalar;
  return TNT_OK;
}

// Stream-ordered variant: the result lands in out[out_index] on the device;
// nothing blocks, so it can be queued behind a matvec and captured.
int tnt_weighted_dot_async(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                           const double* weights, size_t n_regions, tnt_buf* out, size_t out_index) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, a, "a")) return rc;
  if (int rc = check_buf(ctx, b, "b")) return rc;
  if (int rc = check_buf(ctx, out, "out")) return rc;
  if (out_index >= out->len) FAIL(ctx, TNT_E_OOB, "result index outside the output buffer");
  if (n_regions && (!regions || !weights)) FAIL(ctx, TNT_E_INVALID, "null region list");
  std::vector<int64_t> cb, ce;
  std::vector<int32_t> cr;
  std::vector<double> w(weights, weights + n_regions);
  const int64_t chunk = 1 << 16;
  for (size_t r = 0; r < n_regions; ++r) {
    const tnt_region& g = regions[r];
    if (g.offset < 0 || g.len < 0 || (size_t)(g.offset + g.len) > a->len || (size_t)(g.offset + g.len) > b->len)
      FAIL(ctx, TNT_E_OOB, "region outside the payloads");
    for (int64_t s0 = 0; s0 < g.len; s0 += chunk) {
      cb.push_back(g.offset + s0);
      ce.push_back(g.offset + std::min(g.len, s0 + chunk));
      cr.push_back((int32_t)r);
    }
  }
  if (int rc = set_device(ctx)) return rc;
  double* dst = static_cast<double*>(out->ptr) + out_index;
  if (cb.empty()) {
    launch_fill(dst, 1, 0.0, ctx->stream);
    return TNT_OK;
  }
  begin_call(ctx);
  void *dcb, *dce, *dcr, *dw, *parts;
  int rc = cached_table(ctx, cb.data(), 8 * cb.size(), "DB", &dcb);
  if (!rc) rc = cached_table(ctx, ce.data(), 8 * ce.size(), "DE", &dce);
  if (!rc) rc = cached_table(ctx, cr.data(), 4 * cr.size(), "DR", &dcr);
  if (!rc) rc = cached_table(ctx, w.data(), 8 * w.size(), "DW", &dw);
  if (rc) return rc;
  // partials: one grow-only buffer per context; every user is on the context
  // stream, so stream order keeps consecutive calls from overlapping in it
  if (ctx->dot_parts_len < cb.size()) {
    if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "dot scratch must grow during graph capture; run the call once before capturing");
    if (ctx->dot_parts) retire(ctx, ctx->dot_parts);
    ctx->dot_parts = nullptr;
    ctx->dot_parts_len = 0;
    if ((rc = pool_alloc(ctx, 8 * cb.size(), &ctx->dot_parts))) return rc;
    ctx->dot_parts_len = cb.size();
  }
  parts = ctx->dot_parts;
  launch_weighted_dot_partials(static_cast<const double*>(a->ptr), static_cast<const double*>(b->ptr),
                               static_cast<int64_t*>(dcb), static_cast<int64_t*>(dce),
                               static_cast<int32_t*>(dcr), static_cast<double*>(dw),
                               static_cast<double*>(parts), (uint32_t)cb.size(), ctx->stream);
  launch_sum_partials(static_cast<double*>(parts), (uint32_t)cb.size(), dst, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  return TNT_OK;
}

// tnt_weighted_dot_async without any device allocation: chunk tables and
// partials live in this device's `scratch` array (kNoAllocScratchBytes), the
// tables are staged through the context's host buffer.  Same chunking and
// kernels as tnt_weighted_dot_async, so the result is bit-identical.  Every
// partial is written by its chunk's workgroup before it is summed, so
// nothing in the array needs clearing.  Calls on one GPU share the array:
// calls queued on one context are ordered by its stream, calls from two
// contexts (streams) of one GPU that overlap race.
int tnt_weighted_dot_async_noalloc(tnt_ctx* ctx, const tnt_buf* a, const tnt_buf* b, const tnt_region* regions,
                                   const double* weights, size_t n_regions, tnt_buf* out, size_t out_index) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, a, "a")) return rc;
  if (int rc = check_buf(ctx, b, "b")) return rc;
  if (int rc = check_buf(ctx, out, "out")) return rc;
  if (out_index >= out->len) FAIL(ctx, TNT_E_OOB, "result index outside the output buffer");
  if (n_regions && (!regions || !weights)) FAIL(ctx, TNT_E_INVALID, "null region list");
  if (!ctx->scratch_dev) FAIL(ctx, TNT_E_HIP, "allocation-free scratch unavailable: " + ctx->scratch_err);
  const int64_t chunk = 1 << 16;
  size_t n_chunks = 0;
  for (size_t r = 0; r < n_regions; ++r) {
    const tnt_region& g = regions[r];
    if (g.offset < 0 || g.len < 0 || (size_t)(g.offset + g.len) > a->len || (size_t)(g.offset + g.len) > b->len)
      FAIL(ctx, TNT_E_OOB, "region outside the payloads");
    n_chunks += (size_t)((g.len + chunk - 1) / chunk);
  }
  const size_t off_cb = 8 * n_regions, off_ce = off_cb + 8 * n_chunks, off_cr = off_ce + 8 * n_chunks;
  const size_t off_parts = off_cr + ((4 * n_chunks + 7) & ~size_t(7));
  const size_t need = off_parts + 8 * n_chunks;
  if (need > kNoAllocScratchBytes || n_chunks > UINT32_MAX)
    FAIL(ctx, TNT_E_UNSUPPORTED, "region list needs " + std::to_string(need) +
                                     " bytes of scratch; the allocation-free dot holds at most " +
                                     std::to_string(kNoAllocScratchBytes) + " (8 per region + 28 per 65536-element chunk)");
  if (int rc = set_device(ctx)) return rc;
  double* dst = static_cast<double*>(out->ptr) + out_index;
  if (n_chunks == 0) {
    launch_fill(dst, 1, 0.0, ctx->stream);
    return launch_status(ctx);
  }
  unsigned char* hs = ctx->scratch_host.get();
  std::memcpy(hs, weights, 8 * n_regions);
  int64_t* cb = reinterpret_cast<int64_t*>(hs + off_cb);
  int64_t* ce = reinterpret_cast<int64_t*>(hs + off_ce);
  int32_t* cr = reinterpret_cast<int32_t*>(hs + off_cr);
  size_t c = 0;
  for (size_t r = 0; r < n_regions; ++r) {
    const tnt_region& g = regions[r];
    for (int64_t s0 = 0; s0 < g.len; s0 += chunk, ++c) {
      cb[c] = g.offset + s0;
      ce[c] = g.offset + std::min(g.len, s0 + chunk);
      cr[c] = (int32_t)r;
    }
  }
  unsigned char* ds = static_cast<unsigned char*>(ctx->scratch_dev);
  // the tables only; stream order keeps the previous call's kernels ahead of this copy
  HIPCHK(ctx, hipMemcpyAsync(ds, hs, off_parts, hipMemcpyHostToDevice, ctx->stream));
  launch_weighted_dot_partials(static_cast<const double*>(a->ptr), static_cast<const double*>(b->ptr),
                               reinterpret_cast<const int64_t*>(ds + off_cb), reinterpret_cast<const int64_t*>(ds + off_ce),
                               reinterpret_cast<const int32_t*>(ds + off_cr), reinterpret_cast<const double*>(ds),
                               reinterpret_cast<double*>(ds + off_parts), (uint32_t)n_chunks, ctx->stream);
  launch_sum_partials(reinterpret_cast<const double*>(ds + off_parts), (uint32_t)n_chunks, dst, ctx->stream);
  return launch_status(ctx);
}

struct tnt_event {
  hipEvent_t e = nullptr;
};

int tnt_event_create(tnt_ctx* ctx, tnt_event** out) {
  if (!ctx || !out) return TNT_E_INVALID;
  if (int rc = set_device(ctx)) return rc;
  auto* ev = new tnt_event;
  hipError_t e = hipEventCreateWithFlags(&ev->e, hipEventDisableTiming);
  if (e != hipSuccess) {
    delete ev;
    FAIL(ctx, TNT_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
  }
  *out = ev;
  return TNT_OK;
}

int tnt_event_record(tnt_ctx* ctx, tnt_event* ev) {
  if (!ctx || !ev) return TNT_E_INVALID;
  HIPCHK(ctx, hipEventRecord(ev->e, ctx->stream));
  return TNT_OK;
}

// 1: everything queued before the record has finished; 0: still running.
int tnt_event_query(tnt_ctx* ctx, const tnt_event* ev) {
  if (!ctx || !ev) return TNT_E_INVALID;
  hipError_t e = hipEventQuery(ev->e);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  FAIL(ctx, TNT_E_HIP, std::string("hipEventQuery: ") + hipGetErrorString(e));
}

int tnt_event_destroy(tnt_ctx* ctx, tnt_event* ev) {
  if (!ctx || !ev) return TNT_E_INVALID;
  (void)hipEventDestroy(ev->e);
  delete ev;
  return TNT_OK;
}

int tnt_axpby(tnt_ctx* ctx, tnt_buf* y, const tnt_buf* x, double alpha, double beta) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, y, "y")) return rc;
  if (int rc = check_buf(ctx, x, "x")) return rc;
  if (x->len != y->len) FAIL(ctx, TNT_E_SHAPE, "axpby payload lengths differ");
  if (int rc = set_device(ctx)) return rc;
  launch_axpby(static_cast<double*>(y->ptr), static_cast<const double*>(x->ptr), (int64_t)y->len, alpha, beta, ctx->stream);
  HIPCHK(ctx, hipGetLastError());
  return TNT_OK;
}

int tnt_axpby_dev(tnt_ctx* ctx, tnt_buf* y, const tnt_buf* x, const tnt_buf* scalars, int64_t alpha_index,
                  double alpha_factor, int64_t beta_index, double beta_factor, int64_t guard_index) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, y, "y")) return rc;
  if (int rc = check_buf(ctx, x, "x")) return rc;
  if (int rc = check_buf(ctx, scalars, "scalars")) return rc;
  if (x->len != y->len) FAIL(ctx, TNT_E_SHAPE, "axpby payload lengths differ");
  const int64_t ns = (int64_t)scalars->len;
  if (alpha_index >= ns || beta_index >= ns || guard_index >= ns) FAIL(ctx, TNT_E_OOB, "scalar index outside the scalar buffer");
  if (int rc = set_device(ctx)) return rc;
  launch_axpby_dev(static_cast<double*>(y->ptr), static_cast<const double*>(x->ptr), (int64_t)y->len,
                   static_cast<const double*>(scalars->ptr), alpha_index, alpha_factor, beta_index, beta_factor,
                   guard_index, ctx->stream);
  return launch_status(ctx);
}

int tnt_scalars(tnt_ctx* ctx, tnt_buf* scalars, const tnt_scalar_op* ops, size_t n_ops) {
  if (!ctx) return TNT_E_INVALID;
  if (int rc = check_buf(ctx, scalars, "scalars")) return rc;
  if (n_ops > (size_t)kMaxScalarOps) FAIL(ctx, TNT_E_UNSUPPORTED, "at most 8 scalar steps per call");
  if (n_ops && !ops) FAIL(ctx, TNT_E_INVALID, "null scalar program");
  ScalarProgram prog{};
  prog.n = (int32_t)n_ops;
  const int64_t ns = (int64_t)scalars->len;
  for (size_t i = 0; i < n_ops; ++i) {
    const tnt_scalar_op& o = ops[i];
    if (o.op < TNT_SCALAR_SET || o.op > TNT_SCALAR_CG_GUARD) FAIL(ctx, TNT_E_INVALID, "unknown scalar step");
    if (o.out < 0 || o.out >= ns || o.a < 0 || o.a >= ns || o.b < 0 || o.b >= ns)
      FAIL(ctx, TNT_E_OOB, "scalar index outside the scalar buffer");
    prog.ops[i] = ScalarOp{o.op, o.out, o.a, o.b, o.imm};
  }
  if (n_ops == 0) return TNT_OK;
  if (int rc = set_device(ctx)) return rc;
  launch_scalars(static_cast<double*>(scalars->ptr), prog, ctx->stream);
  return launch_status(ctx);
}

int tnt_graph_begin(tnt_ctx* ctx) {
  if (!ctx) return TNT_E_INVALID;
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "graph capture already active");
  if (int rc = set_device(ctx)) return rc;
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  ctx->capture_keys.clear();
  return TNT_OK;
}

int tnt_graph_end(tnt_ctx* ctx, tnt_graph** out) {
  if (!ctx || !out) return TNT_E_INVALID;
  if (!ctx->capturing) FAIL(ctx, TNT_E_INVALID, "no graph capture active");
  ctx->capturing = false;
  auto* g = new tnt_graph;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
  if (e != hipSuccess) {
    delete g;
    ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
    return TNT_E_HIP;
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    delete g;
    ctx->capture_keys.clear();
    ctx->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
    return TNT_E_HIP;
  }
  // pin the cache entries the graph recorded; keep retired memory while it lives
  std::sort(ctx->capture_keys.begin(), ctx->capture_keys.end());
  ctx->capture_keys.erase(std::unique(ctx->capture_keys.begin(), ctx->capture_keys.end()), ctx->capture_keys.end());
  for (const auto& k : ctx->capture_keys) {
    auto it = ctx->tables.find(k);
    if (it != ctx->tables.end()) {
      ++it->second.pins;
      g->pinned.push_back(k);
    }
  }
  ctx->capture_keys.clear();
  ++ctx->live_graphs;
  *out = g;
  return TNT_OK;
}

int tnt_graph_launch(tnt_ctx* ctx, const tnt_graph* g) {
  if (!ctx || !g) return TNT_E_INVALID;
  if (int rc = set_device(ctx)) return rc;
  HIPCHK(ctx, hipGraphLaunch(g->exec, ctx->stream));
  return TNT_OK;
}

int tnt_graph_destroy(tnt_ctx* ctx, tnt_graph* g) {
  if (!ctx || !g) return TNT_E_INVALID;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  for (const auto& k : g->pinned) {
    auto it = ctx->tables.find(k);
    if (it != ctx->tables.end() && it->second.pins > 0) --it->second.pins;
  }
  delete g;
  if (ctx->live_graphs > 0 && --ctx->live_graphs == 0) {
    // nothing can replay the retired memory any more
    for (auto p : ctx->retired) pool_free(ctx, p);
    ctx->retired.clear();
    if (ctx->split) split_workspace_release(ctx->split.get());
  }
  return TNT_OK;
}

int tnt_ctx_set_timing(tnt_ctx* ctx, int on) {
  if (!ctx) return TNT_E_INVALID;
  if (ctx->capturing) FAIL(ctx, TNT_E_INVALID, "timing cannot change during graph capture");
  ctx->timing = on != 0;
  return TNT_OK;
}

int tnt_ctx_kernel_stats(tnt_ctx* ctx, int kind, size_t* launches, double* total_ms, double* total_work) {
  if (!ctx || kind < 0 || kind >= kTimingKinds) return TNT_E_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  size_t n = 0;
  double ms = 0, w = 0;
  for (const auto& t : ctx->timed) {
    if (t.kind != kind) continue;
    float e = 0;
    HIPCHK(ctx, hipEventElapsedTime(&e, t.a, t.b));
    ++n;
    ms += e;
    w += t.work;
  }
  if (launches) *launches = n;
  if (total_ms) *total_ms = ms;
  if (total_work) *total_work = w;
  return TNT_OK;
}

int tnt_ctx_reset_stats(tnt_ctx* ctx) {
  if (!ctx) return TNT_E_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  for (auto& t : ctx->timed) {
    ctx->event_pool.push_back(t.a);
    ctx->event_pool.push_back(t.b);
  }
  ctx->timed.clear();
  return TNT_OK;
}

int tnt_ctx_inject_fault(tnt_ctx* ctx, int kind, int after) {
  if (!ctx) return TNT_E_INVALID;
  if (after < -1) FAIL(ctx, TNT_E_INVALID, "fault countdown must be >= -1");
  switch (kind) {
    case TNT_FAULT_NONE: ctx->fail_alloc_after = ctx->fail_launch_after = -1; return TNT_OK;
    case TNT_FAULT_ALLOC: ctx->fail_alloc_after = after; return TNT_OK;
    case TNT_FAULT_LAUNCH: ctx->fail_launch_after = after; return TNT_OK;
    default: FAIL(ctx, TNT_E_INVALID, "unknown fault kind");
  }
}

int tnt_ctx_set_table_cache_cap(tnt_ctx* ctx, size_t entries) {
  if (!ctx) return TNT_E_INVALID;
  if (entries == 0) FAIL(ctx, TNT_E_INVALID, "table cache cap must be >= 1");
  ctx->table_cap = entries;
  trim_tables(ctx);
  return TNT_OK;
}

// Tile-shape override used by the autotuner sweep (-1 restores the table).
int tnt_set_tile_config(tnt_ctx* ctx, int cfg) {
  if (!ctx) return TNT_E_INVALID;
  if (cfg < -1 || cfg >= TILE_CFG_COUNT) FAIL(ctx, TNT_E_INVALID, "unknown tile configuration");
  ctx->knobs.override_cfg = cfg;
  return TNT_OK;
}

int tnt_set_split_slices(tnt_ctx* ctx, int slices) {
  if (!ctx) return TNT_E_INVALID;
  if (slices < 2 || slices > 6) FAIL(ctx, TNT_E_INVALID, "split slices must be 2..6");
  ctx->split_slices = slices;
  return TNT_OK;
}

int tnt_set_tile_table(tnt_ctx* ctx, const int64_t* m_max, const int64_t* n_max, const int32_t* cfg, size_t n) {
  if (!ctx) return TNT_E_INVALID;
  if (n && (!m_max || !n_max || !cfg)) FAIL(ctx, TNT_E_INVALID, "null tile table");
  std::vector<TileRule> t;
  for (size_t i = 0; i < n; ++i) {
    if (cfg[i] < 0 || cfg[i] >= TILE_CFG_COUNT) FAIL(ctx, TNT_E_INVALID, "unknown tile configuration");
    t.push_back({m_max[i], n_max[i], cfg[i]});
  }
  ctx->knobs.table.swap(t);
  return TNT_OK;
}

}  // extern "C"
