// Split (Ozaki-style) block GEMMs for gfx950: FP64 products from BF16 or INT8
// digit planes.  The BF16 scheme is described here; the INT8 scheme (whole-K
// line exponents, INT8 digits with carries, exact INT32 level sums on
// v_mfma_i32_16x16x64_i8) at split_i8_chunk_exp / split_i8_gemm_kernel below.
// Both share the tables, the workspace and the GEMM stage pipeline.
//
// FP64 C = op(A) op(B) from BF16 matrix-core products, K cut into chunks of
// 256 (one chunk = one exact FP32 window of the GEMM):
//   1. per row i of op(A) and chunk c an exponent e_ic with max_{k in c}
//      |A_ik| < 2^e_ic, per column j of op(B) and chunk c an exponent f_jc;
//   2. a = A_ik 2^-e_ic in (-1, 1) is cut into S base-256 digits by repeated
//      round-to-nearest: d_1 in [-256, 256], d_p in [-128, 128] for p > 1 --
//      small integers, exact in BF16.  Plane p stores d_p 2^-8(p-1) (an exact
//      exponent shift), so a = 2^(e_ic - 8) sum_p plane_p;
//   3. C_ij = sum_c 2^(e_ic + f_jc - 16) sum_{p+q <= S-1 (0-based)} plane_p(A) plane_q(B).
//      Products are summed on v_mfma_f32_16x16x32_bf16 into FP32 accumulator
//      groups by level l = p + q: l = 0 and l = 1 each have their own group;
//      within a K window of 256 every partial sum of those two groups is a
//      multiple of 2^-8l of magnitude at most 2^(24-8l) (|d_1| <= 2^8,
//      |d_p| <= 2^7), so FP32 holds it exactly.
//      All remaining levels share one group (two for S = 6); their FP32
//      rounding is 2^-24 relative to a level-2 sum, i.e. about 2^-40 of the
//      result.  After each window the groups are scaled by the window's
//      2^(e_ic + f_jc - 16) and added into FP64.
// Truncating the digit expansion (relative 2^-8S per operand) and dropping
// the levels l > S-1 are the only other errors: S = 5 gives relative
// Frobenius errors of a few 1e-12 on normal data.
//
// Planes are stored in MFMA fragment order: one fragment = 16 lines x 32 k
// (1 KiB), lane l holding line l%16, k = 8(l/16)..8(l/16)+7, the S planes of
// one (line block, k step) adjacent.  The GEMM copies fragments straight into
// LDS with global_load_lds_dwordx4 (no staging registers), double-buffered in
// two distinct LDS arrays so that the wait for one stage does not wait for the
// prefetch of the next, and every wave reads its operands with one
// conflict-free ds_read_b128 per fragment.  Both operands use the same
// k-to-lane assignment, so the result does not depend on the hardware's
// internal k order; D follows the 16x16 layout (row 4(lane/16)+r, col lane%16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>

#include "../../../include/tenet_hip.h"
#include "split.hpp"
#include "../common/i8_digits.hpp"

namespace tnt_dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

constexpr int FRAG = 512;         // bf16 per fragment (16 lines x 32 k)
constexpr int FRAG_BYTES = 1024;  // one fragment of either scheme: 16 bytes per lane
constexpr int KSTEP = 32;         // k per fragment / MFMA step
constexpr int WINDOW_STEPS = 8;   // 256 k per exact FP32 window

// INT8 scheme: 16 lines x 64 k per fragment, pre-pass items of 512 k, INT32
// level sums exact over windows of 256 steps (16384 k; the bound is 26296).
constexpr int I8_KSTEP = 64;
constexpr int I8_KCHUNK = 512;
constexpr int I8_WINDOW_STEPS = 256;

// The two digit schemes as seen by the GEMM's inner loop: fragment and
// accumulator types, the number of level accumulator groups, the MFMA.
struct Bf16Scheme {
  typedef bf16x8 frag_t;
  typedef f32x4 acc_t;
  static constexpr int nacc(int S) { return S >= 6 ? 4 : 3; }
  static __device__ __forceinline__ acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
struct I8Scheme {
  typedef i32x4 frag_t;
  typedef i32x4 acc_t;
  static constexpr int nacc(int S) { return S; }  // one exact INT32 group per level
  static __device__ __forceinline__ acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
  }
};

struct SplitJob {
  int64_t d_off;     // element offset of C from dbase
  int64_t fa0, fb0;  // first fragment of this job's A / B planes
  int64_t ea0, eb0;  // first exponent of this job's A rows / B columns (chunk-major: ea0 + c * m + i)
  int64_t ra0, rb0;  // INT8 scheme: first whole-K line exponent of this job's A rows / B columns
  int32_t m, n, ns;  // ns = number of 32-k steps
  int32_t ldd;
  int32_t lba, lbb;  // line blocks of 16 (A rows, B columns)
  int32_t cfg;
  int32_t nkc;       // K chunks (= exact windows)
};

// Tile shapes: wave tile (16 WM) x (16 WN), 2 x 2 waves per workgroup.
// 64x64, 128x64, 64x128 with 4 waves; 128x128 with 8 waves (two per SIMD,
// 160 KiB of LDS: measured 81 against 75 TFLOP/s f64-equivalent with 4 waves).
enum SplitShape { SPLIT_64x64 = 0, SPLIT_128x64 = 1, SPLIT_64x128 = 2, SPLIT_128x128 = 3, SPLIT_SHAPES = 4 };
constexpr int shape_rows(int c) { return (c == SPLIT_128x64 || c >= SPLIT_128x128) ? 128 : 64; }
constexpr int shape_cols(int c) { return (c == SPLIT_64x128 || c >= SPLIT_128x128) ? 128 : 64; }

struct SplitTables {
  SplitJob* jobs = nullptr;
  GTile* tiles = nullptr;
  int64_t* prefix = nullptr;   // per-side per-job start of its items in the pre-pass grid, then the
                               // same for the INT8 exponent pass (jobs of one chunk have no items)
  uint32_t tile_start[SPLIT_SHAPES] = {}, tile_count[SPLIT_SHAPES] = {};
  uint32_t items_a = 0, items_b = 0, exp_items_a = 0, exp_items_b = 0;
  int64_t frags_a = 0, frags_b = 0, a_exps = 0, b_exps = 0, a_rows = 0, b_rows = 0;
  uint64_t last_use = 0;  // SplitWorkspace::uses at the table's last run
};

// Job-list tables kept per workspace before least recently used ones are
// dropped (only while no graph of the context is alive).
constexpr size_t kSplitTableCap = 256;

struct SplitWorkspace {
  void* planes = nullptr;  // A planes | B planes (bf16) | A exps | B exps (int32)
  size_t bytes = 0;
  std::unordered_map<std::string, SplitTables> tables;
  uint64_t uses = 0;
  // device memory a live graph may have recorded: outgrown plane buffers and
  // dropped tables, freed by split_workspace_release once no graph is alive
  std::vector<void*> retired;
  static void free_table(SplitTables& t) {
    (void)hipFree(t.jobs);
    (void)hipFree(t.tiles);
    (void)hipFree(t.prefix);
  }
  ~SplitWorkspace() {
    if (planes) (void)hipFree(planes);
    for (auto& kv : tables) free_table(kv.second);
    for (auto p : retired) (void)hipFree(p);
  }
};

SplitWorkspace* split_workspace_create() { return new SplitWorkspace(); }
void split_workspace_destroy(SplitWorkspace* ws) { delete ws; }

void split_workspace_release(SplitWorkspace* ws) {
  if (!ws || ws->retired.empty()) return;
  (void)hipDeviceSynchronize();
  for (auto p : ws->retired) (void)hipFree(p);
  ws->retired.clear();
}

// ---------------------------------------------------------------------------
// Pre-pass: one launch over items (operand side, job, block of 16 lines,
// chunk of 256 k), both operands in one grid.  An item reads its 16 x 256
// elements once into registers, takes the 16 line maxima of the chunk,
// stores their exponents and writes the digit fragments.
//   element (line, kk): side 0 -> op(A)[line][kk], side 1 -> op(B)[kk][line]
// ---------------------------------------------------------------------------
constexpr int KCHUNK = KSTEP * WINDOW_STEPS;  // 256: pre-pass chunk = exact window
constexpr int CHUNK_STEPS = KCHUNK / KSTEP;    // 8 fragments of k per chunk
constexpr int EXP_NONE = -4096;                // exponent of an all-zero line chunk

struct PrepSide {
  uintptr_t base;          // operand base
  int32_t* exps;           // line x chunk exponents of this side
  void* planes;            // fragments of this side
  const int64_t* prefix;   // n_jobs + 1 entries: first item of each job
  int32_t* rows;           // INT8 scheme: whole-K line exponents of this side
};

struct PrepArgs {
  const GJob* jobs;
  const SplitJob* sj;
  int n_jobs;
  uint32_t items_a;
  PrepSide side[2];
};

__device__ __forceinline__ int find_job(const int64_t* prefix, int n, int64_t b) {
  int lo = 0, hi = n;  // prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ int64_t elem_index(const GJob& g, int side, int line, int kk) {
  if (side == 0) return (g.ops & 1) ? (kk + (int64_t)line * g.lda) : (line + (int64_t)kk * g.lda);
  return (g.ops & 2) ? (line + (int64_t)kk * g.ldb) : (kk + (int64_t)line * g.ldb);
}

// Element (line, kk) of the item's operand, 0 outside the job.  Lanes outside
// the job read a zero word instead (the address is selected, not the value),
// so every load is unconditional and its value used: a load whose value only
// feeds a select is sunk into a branch and waited for before the next load.
__device__ double kZeroElem;  // never written: reads as 0
__device__ __forceinline__ double load_elem(const gptr<const double>& M, const GJob& g, int side, int line,
                                            int kk, int lines, int K) {
  const bool ok = line < lines && kk < K;
  const gptr<const double> p =
      ok ? M + elem_index(g, side, line, kk) : (gptr<const double>)(uintptr_t)&kZeroElem;
  return *p;
}

// Element (line, kk) with both indices clamped into the job (lines, K >= 1):
// a duplicate of an element of the same line for kk >= K, of another line for
// line >= lines -- harmless in a line maximum whose padded lines are not stored.
__device__ __forceinline__ double load_clamped(const gptr<const double>& M, const GJob& g, int side, int line,
                                               int kk, int lines, int K) {
  return M[elem_index(g, side, min(line, lines - 1), min(kk, K - 1))];
}

struct PrepItem {
  int side, job, lb, kc;
};

__device__ __forceinline__ PrepItem prep_item(const PrepArgs& x) {
  PrepItem it;
  it.side = blockIdx.x < x.items_a ? 0 : 1;
  const int64_t local = blockIdx.x - (it.side ? x.items_a : 0);
  it.job = find_job(x.side[it.side].prefix, x.n_jobs, local);
  const int nkc = x.sj[it.job].nkc;
  const int64_t r = local - x.side[it.side].prefix[it.job];
  it.lb = (int)(r / nkc);
  it.kc = (int)(r % nkc);
  return it;
}

__device__ __forceinline__ int line_count(const GJob& g, int side) { return side == 0 ? g.m : g.n; }

// Whether consecutive k of one line are adjacent in memory (see elem_index).
// Such sources are read with lanes (lane / 16, lane % 16) on 16 consecutive k
// of lines lane / 16 + 4 r -- four whole 128-byte lines per load -- and staged
// through LDS into fragment order, instead of gathering 8 bytes from 64 lines
// per load.
__device__ __forceinline__ bool k_contiguous(const GJob& g, int side) {
  return side == 0 ? (g.ops & 1) != 0 : (g.ops & 2) == 0;
}

// Column of element (row, col) of a 16 x 64 INT8-step staging tile: the XOR
// keeps both the coalesced writes (4 rows x 16 columns per instruction) and
// the fragment order reads (16 rows x 1 column per 16 lanes) at two LDS passes.
__device__ __forceinline__ int stage_col(int row, int col) { return col ^ (row & 15) ^ ((row & 1) << 4); }

// The same for a 16 x 32 BF16-step tile, read 8 k per lane.
__device__ __forceinline__ int stage_col32(int row, int col) { return col ^ (((row & 1) << 4) | ((row >> 1) & 7)); }

// The S BF16 digit planes of one k step: lane l holds 8 k of line l % 16 (v,
// scaled by the line chunk's exponent e here) and writes 16 bytes of each
// plane frag + p.
template <int S>
__device__ __forceinline__ void bf16_step_digits(double (&v)[8], int e, __bf16* planes, int64_t frag, int lane) {
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = ldexp(v[q], -e);
#pragma unroll
  for (int p = 0; p < S; ++p) {
    // digit d is an integer with |d| <= 256: int -> float -> x 2^-8p -> bf16 is
    // exact at every step (a direct double -> bf16 conversion would be
    // compiled as a round-to-odd sequence of eight instructions)
    const float scale = __builtin_amdgcn_ldexpf(1.0f, -8 * p);
    bf16x8 o;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      v[q] *= 256.0;
      const double d = rint(v[q]);
      v[q] -= d;
      o[q] = (__bf16)((float)(int)d * scale);
    }
    *reinterpret_cast<bf16x8*>(planes + ((frag + p) * FRAG + lane * 8)) = o;
  }
}

// Wave w handles the chunk's k steps w and w + 4; lane l holds line l % 16
// and k = 8 (l / 16) .. + 7 of a step -- the fragment order, so each lane
// writes 16 bytes per plane.  For sources contiguous along lines, the 16 lanes
// of one k read 128 contiguous bytes.  A job whose K fits one k step (ns == 1)
// has one item per four 16-line blocks instead, one block per wave: the same
// exponent and digits without the seven idle steps of a 256-k chunk.
template <int S>
__global__ __launch_bounds__(256) void split_digits(PrepArgs x) {
  static_assert(CHUNK_STEPS == 8, "two k steps per wave");
  __shared__ double red[4][4][16];
  __shared__ int sexp[16];
  __shared__ double stage[4][16][32];  // one 16 x 32 step per wave (its LDS accesses complete in order)
  const PrepItem it = prep_item(x);
  const GJob g = x.jobs[it.job];
  const SplitJob s = x.sj[it.job];
  const PrepSide& sd = x.side[it.side];
  const int lines = it.side == 0 ? g.m : g.n;
  const int K = g.k;
  const gptr<const double> M = gmem<const double>(sd.base) + (it.side == 0 ? g.a_off : g.b_off);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, kg = lane >> 4;
  if (s.ns == 1) {
    const int lb = it.lb * 4 + wave;
    if (lb * 16 >= lines) return;  // whole waves; no workgroup barrier on this path
    const int ln = lb * 16 + li;
    double v[8];
    if (k_contiguous(g, it.side)) {
      double w[4][2];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) w[r][c] = load_elem(M, g, it.side, lb * 16 + kg + 4 * r, 16 * c + li, lines, K);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) stage[wave][kg + 4 * r][stage_col32(kg + 4 * r, 16 * c + li)] = w[r][c];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = stage[wave][li][stage_col32(li, kg * 8 + q)];
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = load_elem(M, g, it.side, ln, kg * 8 + q, lines, K);
    }
    double mx = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) mx = fmax(mx, fabs(v[q]));
    red[wave][kg][li] = mx;
    __builtin_amdgcn_wave_barrier();
    double m2 = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) m2 = fmax(m2, red[wave][c][li]);
    int e = EXP_NONE;
    if (m2 > 0.0) frexp(m2, &e);
    if (e == EXP_NONE) e = 0;
    if (kg == 0 && ln < lines) sd.exps[(it.side == 0 ? s.ea0 : s.eb0) + ln] = e;
    bf16_step_digits<S>(v, e, static_cast<__bf16*>(sd.planes), (it.side == 0 ? s.fa0 : s.fb0) + (int64_t)lb * S, lane);
    return;
  }
  const int line = it.lb * 16 + li;
  double v[2][8];
  if (k_contiguous(g, it.side)) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * CHUNK_STEPS + wave + 4 * h;
      double w[4][2];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c)
          w[r][c] = load_elem(M, g, it.side, it.lb * 16 + kg + 4 * r, st * KSTEP + 16 * c + li, lines, K);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) stage[wave][kg + 4 * r][stage_col32(kg + 4 * r, 16 * c + li)] = w[r][c];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 8; ++q) v[h][q] = stage[wave][li][stage_col32(li, kg * 8 + q)];
      __builtin_amdgcn_wave_barrier();
    }
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * CHUNK_STEPS + wave + 4 * h;
#pragma unroll
      for (int q = 0; q < 8; ++q) v[h][q] = load_elem(M, g, it.side, line, st * KSTEP + kg * 8 + q, lines, K);
    }
  }
  double mx = 0.0;
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int q = 0; q < 8; ++q) mx = fmax(mx, fabs(v[h][q]));
  red[wave][kg][li] = mx;
  __syncthreads();
  if (t < 16) {
    double m2 = 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int c = 0; c < 4; ++c) m2 = fmax(m2, red[w][c][t]);
    int e = EXP_NONE;
    if (m2 > 0.0) frexp(m2, &e);  // m2 = f 2^e, f in [0.5, 1): |x| < 2^e
    if (e == EXP_NONE) e = 0;     // all-zero line chunk: its digits are zero
    sexp[t] = e;
    if (it.lb * 16 + t < lines)
      sd.exps[(it.side == 0 ? s.ea0 : s.eb0) + (int64_t)it.kc * lines + it.lb * 16 + t] = e;
  }
  __syncthreads();
  const int e = sexp[li];
  const int64_t f0 = (it.side == 0 ? s.fa0 : s.fb0) + (int64_t)it.lb * s.ns * S;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int st = it.kc * CHUNK_STEPS + wave + 4 * h;
    if (st >= s.ns) continue;
    bf16_step_digits<S>(v[h], e, static_cast<__bf16*>(sd.planes), f0 + (int64_t)st * S, lane);
  }
}

// ---------------------------------------------------------------------------
// INT8 scheme pre-pass.  Items as above with 512-k chunks; wave w handles the
// chunk's 64-k steps w and w + 4, lane l line l % 16 and k = 16 (l / 16) .. + 15
// of a step (the fragment order: 16 bytes per lane and plane).
//   1. split_i8_chunk_exp: per (line, chunk) the least e with max |A| < 2^e
//      126.5 / 128 (frexp's exponent, plus one for a maximum in the top 1.2 %
//      of its binade), EXP_NONE for an all-zero chunk (chunk-major, as the
//      BF16 exponents);
//   2. split_i8_digits: the line's exponent over all of K, e = max over its
//      chunks (0 for an all-zero line; a job of one chunk skips pass 1 and
//      takes the maximum here), so |a| < 126.5 for a = A 2^(7 - e);
//      a is cut into d_0 = rint(a) in [-126, 126] and S - 1 base-256 digits by
//      repeated round-to-nearest (in [-128, 128]); a digit that reaches 128
//      becomes 128 - 256 and carries one into the digit above, from the last
//      digit up, so every digit fits INT8 (d_0 <= 127) and
//        A = 2^(e - 7) sum_p d_p 2^(-8 p)   (truncated after S digits).
// ---------------------------------------------------------------------------


__global__ __launch_bounds__(256) void split_i8_chunk_exp(PrepArgs x) {
  __shared__ double red[4][1][16];
  const PrepItem it = prep_item(x);
  const SplitJob s = x.sj[it.job];
  if (s.nkc == 1) return;  // (not in the grid) one chunk: split_i8_digits takes the maximum itself
  const GJob g = x.jobs[it.job];
  const PrepSide& sd = x.side[it.side];
  const int lines = line_count(g, it.side);
  const int K = g.k;
  const gptr<const double> M = gmem<const double>(sd.base) + (it.side == 0 ? g.a_off : g.b_off);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, kg = lane >> 4;
  const int line = it.lb * 16 + li;
  if (k_contiguous(g, it.side)) {
    // lanes (lane / 16, lane % 16) read k = 16 c + lane % 16 of lines
    // lane / 16 + 4 r: one load covers four whole 128-byte lines; a line's
    // maximum is then reduced over its 16 lanes
    double m4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * (I8_KCHUNK / I8_KSTEP) + wave + 4 * h;
      double v[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          v[r][c] = load_clamped(M, g, it.side, it.lb * 16 + kg + 4 * r, st * I8_KSTEP + 16 * c + li, lines, K);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) m4[r] = fmax(m4[r], fabs(v[r][c]));
    }
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) m4[r] = fmax(m4[r], __shfl_xor(m4[r], off, 16));
#pragma unroll
    for (int r = 0; r < 4; ++r) red[wave][0][kg + 4 * r] = m4[r];  // all 16 lanes of a group write the same value
  } else {
    double mx = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * (I8_KCHUNK / I8_KSTEP) + wave + 4 * h;
      double v[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = load_clamped(M, g, it.side, line, st * I8_KSTEP + kg * 16 + q, lines, K);
#pragma unroll
      for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(v[q]));
    }
    // fold the four k groups of a line here so that red[w][0][line] holds it
    mx = fmax(mx, __shfl_xor(mx, 16));
    mx = fmax(mx, __shfl_xor(mx, 32));
    if (kg == 0) red[wave][0][li] = mx;
  }
  __syncthreads();
  if (t < 16 && it.lb * 16 + t < lines) {
    double m2 = 0.0;
#pragma unroll
    for (int w = 0; w < 4; ++w) m2 = fmax(m2, red[w][0][t]);
    int e = EXP_NONE;
    if (m2 > 0.0) {
      frexp(m2, &e);                          // m2 = f 2^e, f in [0.5, 1)
      if (ldexp(m2, 7 - e) >= 126.5) ++e;     // keeps rint(A 2^(7-e)) <= 126
    }
    sd.exps[(it.side == 0 ? s.ea0 : s.eb0) + (int64_t)it.kc * lines + it.lb * 16 + t] = e;
  }
}

// The S INT8 digit planes of one 64-k step: lane l holds 16 k of line l % 16
// (v) and writes 16 bytes of each plane frag + p; e is the line's exponent
// over all of K.  a = v 2^(7-e) is cut into d0 = rint(a) and S-1 balanced
// base-256 digits (tnt_digits: for S <= 5 the remainder is rounded once to
// 8 (S-1) bits and its bytes, biased by 128, are the digits; a top digit of
// 128 carries into d0), so every digit fits INT8.  Byte j of plane word w
// is element 4 w + j of the lane's 16.
template <int S>
__device__ __forceinline__ void i8_step_digits(const double (&v)[16], int e, void* planes, int64_t frag, int lane) {
  uint32_t packed[S][4];
  if constexpr (S <= 5) {
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      uint32_t fr[4], lead = 0u;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int d0 = tnt_digits::i8_digits_fixed<S>(ldexp(v[4 * w + j], 7 - e), &fr[j]);  // |a| < 126.5
        lead |= (uint32_t)(d0 & 0xff) << (8 * j);
      }
      packed[0][w] = lead;
      // digit p is byte S - 1 - p of each fr[j]: gather it from the four
      // elements into one word
#pragma unroll
      for (int p = 1; p < S; ++p) {
        const int b = 8 * (S - 1 - p);
        packed[p][w] = ((fr[0] >> b) & 0xffu) | (((fr[1] >> b) & 0xffu) << 8) | (((fr[2] >> b) & 0xffu) << 16) |
                       (((fr[3] >> b) & 0xffu) << 24);
      }
    }
  } else {
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int w = 0; w < 4; ++w) packed[p][w] = 0u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      int d[S];
      tnt_digits::i8_digits_stepwise<S>(ldexp(v[q], 7 - e), d);
#pragma unroll
      for (int p = 0; p < S; ++p) packed[p][q >> 2] |= (uint32_t)(d[p] & 0xff) << (8 * (q & 3));
    }
  }
#pragma unroll
  for (int p = 0; p < S; ++p) {
    const i32x4 o{(int)packed[p][0], (int)packed[p][1], (int)packed[p][2], (int)packed[p][3]};
    *reinterpret_cast<i32x4*>(static_cast<char*>(planes) + (frag + p) * FRAG_BYTES + lane * 16) = o;
  }
}

// One item per (side, job, 16-line block, 512-k chunk), as the BF16 pre-pass;
// a job whose K fits one 64-k step (ns == 1) has one item per four line
// blocks, one block per wave.
template <int S>
__global__ __launch_bounds__(256) void split_i8_digits(PrepArgs x) {
  static_assert(S >= 2 && S <= 6, "digit count");
  __shared__ int sexp[16];
  __shared__ double red[4][4][16];
  __shared__ double stage[4][16][64];  // one 16 x 64 step per wave (its LDS accesses complete in order)
  const PrepItem it = prep_item(x);
  const GJob g = x.jobs[it.job];
  const SplitJob s = x.sj[it.job];
  const PrepSide& sd = x.side[it.side];
  const int lines = line_count(g, it.side);
  const int K = g.k;
  const gptr<const double> M = gmem<const double>(sd.base) + (it.side == 0 ? g.a_off : g.b_off);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int li = lane & 15, kg = lane >> 4;
  if (s.ns == 1) {  // (then nkc == 1 too: the exponent comes from this step)
    const int lb = it.lb * 4 + wave;
    if (lb * 16 >= lines) return;  // whole waves; no workgroup barrier on this path
    const int ln = lb * 16 + li;
    double v[16];
    if (k_contiguous(g, it.side)) {
      double w[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) w[r][c] = load_elem(M, g, it.side, lb * 16 + kg + 4 * r, 16 * c + li, lines, K);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) stage[wave][kg + 4 * r][stage_col(kg + 4 * r, 16 * c + li)] = w[r][c];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = stage[wave][li][stage_col(li, kg * 16 + q)];
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = load_elem(M, g, it.side, ln, kg * 16 + q, lines, K);
    }
    double mx = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(v[q]));
    red[wave][kg][li] = mx;
    __builtin_amdgcn_wave_barrier();
    double m2 = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c) m2 = fmax(m2, red[wave][c][li]);
    int e = 0;  // all-zero line: its digits are zero
    if (m2 > 0.0) {
      frexp(m2, &e);
      if (ldexp(m2, 7 - e) >= 126.5) ++e;
    }
    if (kg == 0 && ln < lines) sd.rows[(it.side == 0 ? s.ra0 : s.rb0) + ln] = e;
    i8_step_digits<S>(v, e, sd.planes, (it.side == 0 ? s.fa0 : s.fb0) + (int64_t)lb * S, lane);
    return;
  }
  const int line = it.lb * 16 + li;
  double v[2][16];
  if (k_contiguous(g, it.side)) {
    // each wave stages its 16 x 64 step through LDS: coalesced loads (lanes
    // (lane / 16, lane % 16) take k = 16 c + lane % 16 of lines lane / 16 +
    // 4 r, four whole 128-byte lines per load), then fragment-order reads.
    // The tile belongs to one wave, whose LDS accesses complete in order.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * (I8_KCHUNK / I8_KSTEP) + wave + 4 * h;
      double w[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
          w[r][c] = load_elem(M, g, it.side, it.lb * 16 + kg + 4 * r, st * I8_KSTEP + 16 * c + li, lines, K);
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) stage[wave][kg + 4 * r][stage_col(kg + 4 * r, 16 * c + li)] = w[r][c];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int q = 0; q < 16; ++q) v[h][q] = stage[wave][li][stage_col(li, kg * 16 + q)];
      __builtin_amdgcn_wave_barrier();
    }
  } else {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int st = it.kc * (I8_KCHUNK / I8_KSTEP) + wave + 4 * h;
#pragma unroll
      for (int q = 0; q < 16; ++q)
        v[h][q] = load_elem(M, g, it.side, line, st * I8_KSTEP + kg * 16 + q, lines, K);
    }
  }
  if (s.nkc == 1) {  // the line's exponent from this chunk (no exponent pass ran)
    double mx = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < 16; ++q) mx = fmax(mx, fabs(v[h][q]));
    red[wave][kg][li] = mx;
    __syncthreads();
    if (t < 16) {
      double m2 = 0.0;
#pragma unroll
      for (int w = 0; w < 4; ++w)
#pragma unroll
        for (int c = 0; c < 4; ++c) m2 = fmax(m2, red[w][c][t]);
      int e = 0;  // all-zero line: its digits are zero
      if (m2 > 0.0) {
        frexp(m2, &e);
        if (ldexp(m2, 7 - e) >= 126.5) ++e;
      }
      sexp[t] = e;
      const int l = it.lb * 16 + t;
      if (l < lines) sd.rows[(it.side == 0 ? s.ra0 : s.rb0) + l] = e;
    }
  } else if (t < 16) {
    const int l = it.lb * 16 + t;
    int e = EXP_NONE;
    if (l < lines) {
      const int32_t* ce = sd.exps + (it.side == 0 ? s.ea0 : s.eb0) + l;
      for (int c = 0; c < s.nkc; ++c) e = max(e, ce[(int64_t)c * lines]);
    }
    if (e == EXP_NONE) e = 0;  // all-zero line: its digits are zero
    sexp[t] = e;
    if (it.kc == 0 && l < lines) sd.rows[(it.side == 0 ? s.ra0 : s.rb0) + l] = e;
  }
  __syncthreads();
  const int e = sexp[li];
  const int64_t f0 = (it.side == 0 ? s.fa0 : s.fb0) + (int64_t)it.lb * s.ns * S;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int st = it.kc * (I8_KCHUNK / I8_KSTEP) + wave + 4 * h;
    if (st >= s.ns) continue;
    i8_step_digits<S>(v[h], e, sd.planes, f0 + (int64_t)st * S, lane);
  }
}

// ---------------------------------------------------------------------------
// Split GEMM kernel
// ---------------------------------------------------------------------------
struct SplitGemmArgs {
  const SplitJob* sj;
  const GTile* tiles;
  uintptr_t dbase;
  const void* a_planes;
  const void* b_planes;
  const int32_t* a_exp;  // BF16: line x chunk exponents; INT8: whole-K line exponents
  const int32_t* b_exp;
  double alpha, beta;
};

__device__ __forceinline__ void wg_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Workgroup b of a launch runs on XCD b % 8 (workgroups are dispatched to the
// XCDs round-robin).  Tile lists are padded to a multiple of 256 and read in
// super-blocks of 32 consecutive tiles, super-block s on XCD s % 8, so the
// workgroups one XCD runs side by side cover one block of 4 x 8 tiles of a
// job: their k steps stream 4 A and 8 B panels through that XCD's L2 instead
// of one A and 32 B panels.  Bijective for b < 256 n.
constexpr uint32_t kSuperBlock = 32, kXcds = 8;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b) {
  const uint32_t xcd = b % kXcds, local = b / kXcds;
  return ((local / kSuperBlock) * kXcds + xcd) * kSuperBlock + local % kSuperBlock;
}

// 16 bytes per lane from global memory straight into LDS at lds + 16 * lane:
// global address = base (uniform) + voff (per lane).  The copy has no
// register result, so nothing is left in flight in a VGPR; the kernel waits
// for the copies with explicit vmcnt before the barrier that publishes them.
__device__ __forceinline__ void copy_to_lds(uint64_t base, uint32_t voff, uint32_t lds) {
  // M0 holds the LDS destination of the copy.  The asm saves the compiler's
  // M0 and restores it after issuing the copy (which latches M0 at issue, as
  // the compiler's own code for __builtin_amdgcn_global_load_lds assumes), so
  // no compiler-held M0 value is changed; "memory" keeps the copy where the
  // stage pipeline issues it.  (The builtin form schedules the loop
  // differently and ran the BF16 split GEMM 4 % slower.)
  uint32_t saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\ts_mov_b32 m0, %3\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, %2\n\ts_mov_b32 m0, %0"
      : "=&s"(saved)
      : "v"(voff), "s"(base), "s"(lds)
      : "memory");
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t lds_address(const void* p) {
  return __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(const __attribute__((address_space(3))) char*)p);
}

// s_waitcnt with only vmcnt constrained (expcnt and lgkmcnt at their maxima).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup = GM x GN waves (wave w at row w % GM, column w / GM), each
// owning (16 WM) x (16 WN) of the output tile.
template <class Sch, int S, int WM, int WN, int GM = 2, int GN = 2>
struct SplitKernel {
  typedef typename Sch::frag_t frag_t;
  typedef typename Sch::acc_t acc_t;
  static constexpr int NACC = Sch::nacc(S);
  static constexpr int NW = GM * GN;                   // waves per workgroup
  static constexpr int LBM = GM * WM, LBN = GN * WN;   // line blocks per workgroup tile
  static constexpr int FA = LBM * S, FB = LBN * S;     // fragments per stage
  static constexpr int PER_WAVE = (FA + FB) / NW;
  static_assert(FA % NW == 0 && FB % NW == 0, "fragments per stage must split evenly over the waves");

  // Step-invariant part of this wave's copies: fragment r of every stage
  // comes from src[r] + step * S KiB and goes to LDS slot f = wave + NW r of
  // the stage buffer (A fragments for r < FA/NW, then B).
  struct Copies {
    uint64_t src[PER_WAVE];
  };

  static __device__ __forceinline__ void setup(const SplitGemmArgs& g, const SplitJob& job, int tm, int tn,
                                               int wave, Copies& c) {
#pragma unroll
    for (int r = 0; r < PER_WAVE; ++r) {
      const int f = wave + NW * r;
      int64_t frag;
      uintptr_t planes;
      if (r < FA / NW) {
        const int i = f / S, p = f % S;
        const int lb = min(tm * LBM + i, job.lba - 1);  // clamp: rows past m are discarded
        frag = job.fa0 + (int64_t)lb * job.ns * S + p;
        planes = (uintptr_t)g.a_planes;
      } else {
        const int fb = f - FA;
        const int jb = fb / S, q = fb % S;
        const int lb = min(tn * LBN + jb, job.lbb - 1);
        frag = job.fb0 + (int64_t)lb * job.ns * S + q;
        planes = (uintptr_t)g.b_planes;
      }
      c.src[r] = uniform64(planes + (uint64_t)frag * FRAG_BYTES);
    }
  }

  // Issue this wave's share of one stage's fragments (always PER_WAVE copies)
  // into the stage buffers at LDS addresses as_lds (A) and bs_lds (B).
  static __device__ __forceinline__ void load_stage(const Copies& c, int step, uint32_t as_lds, uint32_t bs_lds,
                                                    int wave, uint32_t voff) {
    const uint64_t step_bytes = (uint64_t)step * (S * FRAG_BYTES);
#pragma unroll
    for (int r = 0; r < PER_WAVE; ++r) {
      const int f = wave + NW * r;
      const uint32_t lds = r < FA / NW ? as_lds + f * FRAG_BYTES : bs_lds + (f - FA) * FRAG_BYTES;
      copy_to_lds(c.src[r] + step_bytes, voff, lds);
    }
  }

  static __device__ __forceinline__ void compute(frag_t (&As)[FA][64], frag_t (&Bs)[FB][64], int wm, int wn,
                                                 int lane, acc_t (&acc)[NACC][WM][WN]) {
    frag_t bq[S][WN];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < WN; ++j) bq[q][j] = Bs[(wn * WN + j) * S + q][lane];
#pragma unroll
    for (int p = 0; p < S; ++p) {
      frag_t ap[WM];
#pragma unroll
      for (int i = 0; i < WM; ++i) ap[i] = As[(wm * WM + i) * S + p][lane];
#pragma unroll
      for (int q = 0; q < S - p; ++q) {
        constexpr int last = NACC - 1;
        const int grp = (p + q) < last ? (p + q) : last;
#pragma unroll
        for (int i = 0; i < WM; ++i)
#pragma unroll
          for (int j = 0; j < WN; ++j)
            acc[grp][i][j] = Sch::mfma(ap[i], bq[q][j], acc[grp][i][j]);
      }
    }
  }

  static __device__ __forceinline__ void zero(acc_t (&acc)[NACC][WM][WN]) {
#pragma unroll
    for (int l = 0; l < NACC; ++l)
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[l][i][j] = acc_t{0, 0, 0, 0};
  }

  // acc64 += 2^(e_row + f_col - 16) (sum of the window's groups); ea/eb are the
  // window's exponents of this lane's rows and columns.
  static __device__ __forceinline__ void flush(acc_t (&acc)[NACC][WM][WN], double (&acc64)[WM][WN][4],
                                               const int (&ea)[WM][4], const int (&eb)[WN]) {
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double s = 0.0;
#pragma unroll
          for (int l = NACC - 1; l >= 0; --l) s += (double)acc[l][i][j][r];
          acc64[i][j][r] = fma(s, ldexp(1.0, ea[i][r] + eb[j] - 16), acc64[i][j][r]);
        }
  }

  // One k step: prefetch the next stage into (An, Bn), wait for (Ac, Bc), compute.
  static __device__ __forceinline__ void step(const Copies& c, int ns, int s,
                                              frag_t (&Ac)[FA][64], frag_t (&Bc)[FB][64],
                                              uint32_t an_lds, uint32_t bn_lds, int wave, int lane,
                                              acc_t (&acc)[NACC][WM][WN]) {
    if (s + 1 < ns) {
      load_stage(c, s + 1, an_lds, bn_lds, wave, lane * 16);
      wait_vmcnt<PER_WAVE>();
    } else {
      wait_vmcnt<0>();
    }
    wg_barrier();
    compute(Ac, Bc, wave % GM, wave / GM, lane, acc);
    wg_barrier();  // every wave is done with (Ac, Bc) before they are refilled
  }
};

template <int S, int WM, int WN, int GM = 2, int GN = 2>
__global__ __launch_bounds__(64 * GM * GN) void split_gemm_kernel(SplitGemmArgs g) {
  using K = SplitKernel<Bf16Scheme, S, WM, WN, GM, GN>;
  __shared__ bf16x8 A0[K::FA][64];
  __shared__ bf16x8 B0[K::FB][64];
  __shared__ bf16x8 A1[K::FA][64];
  __shared__ bf16x8 B1[K::FB][64];
  const GTile tile = g.tiles[xcd_tile(blockIdx.x)];
  if (tile.job < 0) return;  // padding
  const SplitJob job = g.sj[tile.job];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave % GM, wn = wave / GM;

  f32x4 acc[K::NACC][WM][WN];
  double acc64[WM][WN][4];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[i][j][r] = 0.0;
  typename K::Copies cp;
  K::setup(g, job, tile.tm, tile.tn, wave, cp);
  const int ns = job.ns;
  const uint32_t a0 = lds_address(&A0[0][0]), b0 = lds_address(&B0[0][0]);
  const uint32_t a1 = lds_address(&A1[0][0]), b1 = lds_address(&B1[0][0]);
  K::load_stage(cp, 0, a0, b0, wave, lane * 16);
  const int m0 = tile.tm * K::LBM * 16 + wm * WM * 16, n0 = tile.tn * K::LBN * 16 + wn * WN * 16;
  // windows of WINDOW_STEPS (even) steps = pre-pass chunks: every window starts
  // on stage buffer 0 and has its own row / column exponents
  for (int w0 = 0, kc = 0; w0 < ns; w0 += WINDOW_STEPS, ++kc) {
    const int w1 = min(ns, w0 + WINDOW_STEPS);
    K::zero(acc);
    for (int s = w0; s < w1; s += 2) {
      K::step(cp, ns, s, A0, B0, a1, b1, wave, lane, acc);
      if (s + 1 < w1) K::step(cp, ns, s + 1, A1, B1, a0, b0, wave, lane, acc);
    }
    int ea[WM][4], eb[WN];
    const int32_t* ra = g.a_exp + job.ea0 + (int64_t)kc * job.m;
    const int32_t* rb = g.b_exp + job.eb0 + (int64_t)kc * job.n;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + i * 16 + (lane >> 4) * 4 + r;
        ea[i][r] = row < job.m ? ra[row] : 0;
      }
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int col = n0 + j * 16 + (lane & 15);
      eb[j] = col < job.n ? rb[col] : 0;
    }
    K::flush(acc, acc64, ea, eb);
  }

  // epilogue: C = alpha acc64 (+ beta C)
  const gptr<double> D = gmem<double>(g.dbase) + job.d_off;
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int col = n0 + j * 16 + (lane & 15);
    if (col >= job.n) continue;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + i * 16 + (lane >> 4) * 4 + r;
        if (row < job.m) {
          const double v = g.alpha * acc64[i][j][r];
          const gptr<double> p = D + row + (int64_t)col * job.ldd;
          *p = (g.beta != 0.0) ? v + g.beta * *p : v;
        }
      }
    }
  }
}

// INT8 scheme GEMM: the stage pipeline of split_gemm_kernel over 64-k INT8
// fragments (v_mfma_i32_16x16x64_i8).  Level l = p + q of the digit products
// has its own INT32 accumulator group; over a window of at most 16384 k every
// partial sum is bounded by 81664 k < 2^31 (|d_0| <= 127, |d_p| <= 128), so
// the sums are exact.  At the end of a window
//   C_ij += 2^(e_i + f_j - 14) sum_l 2^(-8 l) L_l
// in FP64 (e_i, f_j: whole-K line exponents); a window after the first adds
// into the D elements that the same lane wrote for the first.
template <int S, int WM, int WN, int GM = 2, int GN = 2>
__global__ __launch_bounds__(64 * GM * GN) void split_i8_gemm_kernel(SplitGemmArgs g) {
  using K = SplitKernel<I8Scheme, S, WM, WN, GM, GN>;
  __shared__ i32x4 A0[K::FA][64];
  __shared__ i32x4 B0[K::FB][64];
  __shared__ i32x4 A1[K::FA][64];
  __shared__ i32x4 B1[K::FB][64];
  const GTile tile = g.tiles[xcd_tile(blockIdx.x)];
  if (tile.job < 0) return;  // padding
  const SplitJob job = g.sj[tile.job];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave % GM, wn = wave / GM;

  i32x4 acc[S][WM][WN];
  typename K::Copies cp;
  K::setup(g, job, tile.tm, tile.tn, wave, cp);
  const int ns = job.ns;
  const uint32_t a0 = lds_address(&A0[0][0]), b0 = lds_address(&B0[0][0]);
  const uint32_t a1 = lds_address(&A1[0][0]), b1 = lds_address(&B1[0][0]);
  K::load_stage(cp, 0, a0, b0, wave, lane * 16);
  const int m0 = tile.tm * K::LBM * 16 + wm * WM * 16, n0 = tile.tn * K::LBN * 16 + wn * WN * 16;
  const gptr<double> D = gmem<double>(g.dbase) + job.d_off;
  for (int w0 = 0; w0 < ns; w0 += I8_WINDOW_STEPS) {  // even window length: each starts on buffer 0
    const int w1 = min(ns, w0 + I8_WINDOW_STEPS);
    K::zero(acc);
    for (int s = w0; s < w1; s += 2) {
      K::step(cp, ns, s, A0, B0, a1, b1, wave, lane, acc);
      if (s + 1 < w1) K::step(cp, ns, s + 1, A1, B1, a0, b0, wave, lane, acc);
    }
    // line exponents (indices clamped into the job; rows and columns past it
    // are not stored)
    int ea[WM][4], eb[WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) ea[i][r] = g.a_exp[job.ra0 + min(m0 + i * 16 + (lane >> 4) * 4 + r, job.m - 1)];
#pragma unroll
    for (int j = 0; j < WN; ++j) eb[j] = g.b_exp[job.rb0 + min(n0 + j * 16 + (lane & 15), job.n - 1)];
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int col = n0 + j * 16 + (lane & 15);
      if (col >= job.n) continue;
#pragma unroll
      for (int i = 0; i < WM; ++i) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + i * 16 + (lane >> 4) * 4 + r;
          if (row < job.m) {
            double sum = 0.0;
#pragma unroll
            for (int l = S - 1; l >= 0; --l) sum += ldexp((double)acc[l][i][j][r], -8 * l);
            const double v = g.alpha * ldexp(sum, ea[i][r] + eb[j] - 14);
            const gptr<double> p = D + row + (int64_t)col * job.ldd;
            if (w0 == 0) *p = (g.beta != 0.0) ? v + g.beta * *p : v;
            else *p += v;
          }
        }
      }
    }
  }
}

template <int S>
static void launch_split_s(const SplitTables& t, const SplitGemmArgs& a0, hipStream_t s) {
  SplitGemmArgs a = a0;
  for (int c = 0; c < SPLIT_SHAPES; ++c) {
    if (!t.tile_count[c]) continue;
    a.tiles = t.tiles + t.tile_start[c];
    const dim3 grid(t.tile_count[c]);
    if constexpr (S >= 6) {  // the wide shapes would spill at six planes
      hipLaunchKernelGGL((split_gemm_kernel<S, 2, 2>), grid, dim3(256), 0, s, a);
    } else {
      if (c == SPLIT_64x64) hipLaunchKernelGGL((split_gemm_kernel<S, 2, 2>), grid, dim3(256), 0, s, a);
      else if (c == SPLIT_128x64) hipLaunchKernelGGL((split_gemm_kernel<S, 4, 2>), grid, dim3(256), 0, s, a);
      else if (c == SPLIT_64x128) hipLaunchKernelGGL((split_gemm_kernel<S, 2, 4>), grid, dim3(256), 0, s, a);
      else if constexpr (S <= 5) hipLaunchKernelGGL((split_gemm_kernel<S, 4, 2, 2, 4>), grid, dim3(512), 0, s, a);
    }
  }
}

template <int S>
static void launch_digits_s(const PrepArgs& pa, uint32_t items, hipStream_t s) {
  if (items) hipLaunchKernelGGL(split_digits<S>, dim3(items), dim3(256), 0, s, pa);
}

template <int S>
static void launch_split_i8_s(const SplitTables& t, const SplitGemmArgs& a0, hipStream_t s) {
  SplitGemmArgs a = a0;
  for (int c = 0; c < SPLIT_SHAPES; ++c) {
    if (!t.tile_count[c]) continue;
    a.tiles = t.tiles + t.tile_start[c];
    const dim3 grid(t.tile_count[c]);
    if constexpr (S >= 6) {  // the wide shapes would spill at six planes
      hipLaunchKernelGGL((split_i8_gemm_kernel<S, 2, 2>), grid, dim3(256), 0, s, a);
    } else {
      if (c == SPLIT_64x64) hipLaunchKernelGGL((split_i8_gemm_kernel<S, 2, 2>), grid, dim3(256), 0, s, a);
      else if (c == SPLIT_128x64) hipLaunchKernelGGL((split_i8_gemm_kernel<S, 4, 2>), grid, dim3(256), 0, s, a);
      else if (c == SPLIT_64x128) hipLaunchKernelGGL((split_i8_gemm_kernel<S, 2, 4>), grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((split_i8_gemm_kernel<S, 4, 2, 2, 4>), grid, dim3(512), 0, s, a);
    }
  }
}

template <int S>
static void launch_digits_i8_s(const PrepArgs& pa, uint32_t items, hipStream_t s) {
  if (items) hipLaunchKernelGGL(split_i8_digits<S>, dim3(items), dim3(256), 0, s, pa);
}

#define SPLIT_HIP(expr)                                                    \
  do {                                                                     \
    hipError_t _e = (expr);                                                \
    if (_e != hipSuccess) {                                                \
      err = std::string(#expr) + ": " + hipGetErrorString(_e);             \
      return _e == hipErrorOutOfMemory ? TNT_E_OOM : TNT_E_HIP;            \
    }                                                                      \
  } while (0)

// Tile shape of one job.  The 128-line shapes hold one workgroup per CU
// (LDS) against two for 64x64; they win only when a tile runs many k steps
// (measured: K ~ 7000 +12%, K ~ 400 -15%), and only where the job fills them.
static int split_shape_for(const GJob& j, int64_t total_64x64_tiles, int S) {
  if (S >= 6 || total_64x64_tiles < 512 || j.k < 1024) return SPLIT_64x64;  // 128x128 at K >= 256: -12 % (U(1) env)
  if (j.m >= 256 && j.n >= 256) return SPLIT_128x128;
  if (j.m >= 128 && j.m >= j.n) return SPLIT_128x64;
  if (j.n >= 128) return SPLIT_64x128;
  return SPLIT_64x64;
}

int split_gemm_run(SplitWorkspace* ws, int scheme, int S, const std::vector<GJob>& host_jobs, const GJob* dev_jobs,
                   uintptr_t d, uintptr_t a, uintptr_t b, double alpha, double beta, hipStream_t s, bool capturing,
                   bool graphs_alive, std::string& err, hipStream_t side, hipEvent_t side_done, int side_op) {
  if (S < 2 || S > 6) {
    err = "split slices outside 2..6";
    return TNT_E_INVALID;
  }
  if (host_jobs.empty()) return TNT_OK;
  const bool i8 = scheme == SPLIT_SCHEME_I8;
  const int kstep = i8 ? I8_KSTEP : KSTEP, kchunk = i8 ? I8_KCHUNK : KCHUNK;
  std::string key(reinterpret_cast<const char*>(host_jobs.data()), host_jobs.size() * sizeof(GJob));
  key.push_back((char)S);
  key.push_back((char)scheme);
  auto it = ws->tables.find(key);
  if (it == ws->tables.end()) {
    if (capturing) {
      err = "split GEMM tables are built on first use; run the call once before capturing";
      return TNT_E_INVALID;
    }
    if (ws->tables.size() >= kSplitTableCap) {
      // drop the least recently used half; memory a live graph may hold is retired
      std::vector<std::pair<uint64_t, std::string>> byuse;
      for (auto& kv : ws->tables) byuse.emplace_back(kv.second.last_use, kv.first);
      std::sort(byuse.begin(), byuse.end());
      SPLIT_HIP(hipStreamSynchronize(s));
      for (size_t i = 0; i < byuse.size() / 2; ++i) {
        auto e = ws->tables.find(byuse[i].second);
        if (graphs_alive) {
          ws->retired.push_back(e->second.jobs);
          ws->retired.push_back(e->second.tiles);
          ws->retired.push_back(e->second.prefix);
        } else {
          SplitWorkspace::free_table(e->second);
        }
        ws->tables.erase(e);
      }
    }
    SplitTables t;
    std::vector<SplitJob> sj(host_jobs.size());
    std::vector<GTile> per[SPLIT_SHAPES];
    std::vector<int64_t> pre_a{0}, pre_b{0}, xpre_a{0}, xpre_b{0};
    int64_t small_tiles = 0;
    for (const GJob& j : host_jobs) small_tiles += (int64_t)((j.m + 63) / 64) * ((j.n + 63) / 64);
    for (size_t i = 0; i < host_jobs.size(); ++i) {
      const GJob& j = host_jobs[i];
      SplitJob x{};
      x.d_off = j.d_off;
      x.m = j.m;
      x.n = j.n;
      x.ns = (std::max(j.k, 1) + kstep - 1) / kstep;
      x.ldd = j.ldd;
      x.lba = (j.m + 15) / 16;
      x.lbb = (j.n + 15) / 16;
      x.fa0 = t.frags_a;
      x.fb0 = t.frags_b;
      t.frags_a += (int64_t)x.lba * x.ns * S;
      t.frags_b += (int64_t)x.lbb * x.ns * S;
      x.nkc = (std::max(j.k, 1) + kchunk - 1) / kchunk;
      x.ea0 = t.a_exps;
      x.eb0 = t.b_exps;
      t.a_exps += (int64_t)j.m * x.nkc;
      t.b_exps += (int64_t)j.n * x.nkc;
      x.ra0 = t.a_rows;
      x.rb0 = t.b_rows;
      if (i8) {
        t.a_rows += j.m;
        t.b_rows += j.n;
      }
      x.cfg = (j.m > 0 && j.n > 0) ? split_shape_for(j, small_tiles, S) : SPLIT_64x64;
      // operands of a job with an empty output are never read: no pre-pass items
      // one-step jobs (ns == 1): one item per four 16-line blocks (one per wave)
      const bool one_step = x.ns == 1;
      const int64_t ia = (j.n > 0) ? (one_step ? (x.lba + 3) / 4 : (int64_t)x.lba * x.nkc) : 0;
      const int64_t ib = (j.m > 0) ? (one_step ? (x.lbb + 3) / 4 : (int64_t)x.lbb * x.nkc) : 0;
      sj[i] = x;
      pre_a.push_back(pre_a.back() + ia);
      pre_b.push_back(pre_b.back() + ib);
      xpre_a.push_back(xpre_a.back() + (x.nkc > 1 ? ia : 0));
      xpre_b.push_back(xpre_b.back() + (x.nkc > 1 ? ib : 0));
      if (j.m == 0 || j.n == 0) continue;
      // blocks of 4 tile rows x 8 tile columns, consecutive (see xcd_tile)
      const int bm = shape_rows(x.cfg), bn = shape_cols(x.cfg);
      const int ntm = (j.m + bm - 1) / bm, ntn = (j.n + bn - 1) / bn;
      for (int r0 = 0; r0 < ntm; r0 += 4)
        for (int c0 = 0; c0 < ntn; c0 += 8)
          for (int tm = r0; tm < std::min(ntm, r0 + 4); ++tm)
            for (int tn = c0; tn < std::min(ntn, c0 + 8); ++tn) per[x.cfg].push_back({(int32_t)i, tm, tn, x.cfg});
    }
    std::vector<GTile> tiles;
    for (int c = 0; c < SPLIT_SHAPES; ++c) {
      if (!per[c].empty())  // pad to whole rounds of super-blocks over the XCDs (no-op tiles)
        per[c].resize((per[c].size() + kSuperBlock * kXcds - 1) / (kSuperBlock * kXcds) * (kSuperBlock * kXcds),
                      GTile{-1, 0, 0, c});
      t.tile_start[c] = (uint32_t)tiles.size();
      t.tile_count[c] = (uint32_t)per[c].size();
      tiles.insert(tiles.end(), per[c].begin(), per[c].end());
    }
    if (pre_a.back() + pre_b.back() > (int64_t)UINT32_MAX) {
      err = "split pre-pass grid too large";
      return TNT_E_UNSUPPORTED;
    }
    t.items_a = (uint32_t)pre_a.back();
    t.items_b = (uint32_t)pre_b.back();
    t.exp_items_a = (uint32_t)xpre_a.back();
    t.exp_items_b = (uint32_t)xpre_b.back();
    std::vector<int64_t> pre;
    pre.insert(pre.end(), pre_a.begin(), pre_a.end());
    pre.insert(pre.end(), pre_b.begin(), pre_b.end());
    pre.insert(pre.end(), xpre_a.begin(), xpre_a.end());
    pre.insert(pre.end(), xpre_b.begin(), xpre_b.end());
    void *pj = nullptr, *pt = nullptr, *pp = nullptr;
    hipError_t e1 = hipMalloc(&pj, sizeof(SplitJob) * sj.size());
    hipError_t e2 = e1 == hipSuccess ? hipMalloc(&pt, sizeof(GTile) * std::max<size_t>(tiles.size(), 1)) : e1;
    hipError_t e3 = e2 == hipSuccess ? hipMalloc(&pp, sizeof(int64_t) * pre.size()) : e2;
    if (e3 != hipSuccess) {
      (void)hipFree(pj);
      (void)hipFree(pt);
      (void)hipGetLastError();
      err = std::string("hipMalloc (split GEMM tables): ") + hipGetErrorString(e3);
      return e3 == hipErrorOutOfMemory ? TNT_E_OOM : TNT_E_HIP;
    }
    t.jobs = static_cast<SplitJob*>(pj);
    t.tiles = static_cast<GTile*>(pt);
    t.prefix = static_cast<int64_t*>(pp);
    it = ws->tables.emplace(key, t).first;
    SPLIT_HIP(hipMemcpyAsync(pj, sj.data(), sizeof(SplitJob) * sj.size(), hipMemcpyHostToDevice, s));
    if (!tiles.empty()) SPLIT_HIP(hipMemcpyAsync(pt, tiles.data(), sizeof(GTile) * tiles.size(), hipMemcpyHostToDevice, s));
    SPLIT_HIP(hipMemcpyAsync(pp, pre.data(), sizeof(int64_t) * pre.size(), hipMemcpyHostToDevice, s));
    SPLIT_HIP(hipStreamSynchronize(s));
  }
  it->second.last_use = ++ws->uses;
  const SplitTables& t = it->second;
  // workspace: A fragments | B fragments (1 KiB each) | A exps | B exps (int32,
  // line x chunk) | A line exps | B line exps (int32, INT8 scheme)
  const size_t bytes = (size_t)FRAG_BYTES * (size_t)(t.frags_a + t.frags_b) +
                       4 * (size_t)(t.a_exps + t.b_exps + t.a_rows + t.b_rows) + 256;
  if (bytes > ws->bytes) {
    if (capturing) {
      err = "split GEMM workspace must be sized before graph capture; run the call once first";
      return TNT_E_INVALID;
    }
    SPLIT_HIP(hipStreamSynchronize(s));
    if (ws->planes) {
      if (graphs_alive) ws->retired.push_back(ws->planes);  // a graph may have recorded it
      else SPLIT_HIP(hipFree(ws->planes));
    }
    ws->planes = nullptr;
    ws->bytes = 0;
    SPLIT_HIP(hipMalloc(&ws->planes, bytes));
    ws->bytes = bytes;
  }
  char* ap = static_cast<char*>(ws->planes);
  char* bp = ap + (size_t)FRAG_BYTES * t.frags_a;
  int32_t* ae = reinterpret_cast<int32_t*>(bp + (size_t)FRAG_BYTES * t.frags_b);
  int32_t* be = ae + t.a_exps;
  int32_t* ar = be + t.b_exps;
  int32_t* br = ar + t.a_rows;
  const int n_jobs = (int)host_jobs.size();
  PrepArgs pa{dev_jobs, t.jobs, n_jobs, t.items_a,
              {{a, ae, ap, t.prefix, ar}, {b, be, bp, t.prefix + (n_jobs + 1), br}}};
  const uint32_t items = t.items_a + t.items_b;
  if (i8) {
    PrepArgs px = pa;  // exponent pass: items of the jobs of more than one chunk
    px.items_a = t.exp_items_a;
    px.side[0].prefix = t.prefix + 2 * (n_jobs + 1);
    px.side[1].prefix = t.prefix + 3 * (n_jobs + 1);
    auto digits = [&](const PrepArgs& q, uint32_t n, hipStream_t st) {
      switch (S) {
        case 2: launch_digits_i8_s<2>(q, n, st); break;
        case 3: launch_digits_i8_s<3>(q, n, st); break;
        case 4: launch_digits_i8_s<4>(q, n, st); break;
        case 6: launch_digits_i8_s<6>(q, n, st); break;
        default: launch_digits_i8_s<5>(q, n, st); break;
      }
    };
    if (side) {
      // operand b's pre-passes as grids of its items only (items_a = 0),
      // operand a's as grids of the first items_a items; operand side_op's on
      // `side`, the other's here; joined before the GEMM
      PrepArgs pxb = px, pab = pa;
      pxb.items_a = 0;
      pab.items_a = 0;
      hipStream_t sa = side_op == 0 ? side : s, sb = side_op == 0 ? s : side;
      if (t.exp_items_b) hipLaunchKernelGGL(split_i8_chunk_exp, dim3(t.exp_items_b), dim3(256), 0, sb, pxb);
      digits(pab, t.items_b, sb);
      if (t.exp_items_a) hipLaunchKernelGGL(split_i8_chunk_exp, dim3(t.exp_items_a), dim3(256), 0, sa, px);
      digits(pa, t.items_a, sa);
      SPLIT_HIP(hipEventRecord(side_done, side));
      SPLIT_HIP(hipStreamWaitEvent(s, side_done, 0));
    } else {
      if (t.exp_items_a + t.exp_items_b)
        hipLaunchKernelGGL(split_i8_chunk_exp, dim3(t.exp_items_a + t.exp_items_b), dim3(256), 0, s, px);
      digits(pa, items, s);
    }
    SplitGemmArgs g{t.jobs, t.tiles, d, ap, bp, ar, br, alpha, beta};
    switch (S) {
      case 2: launch_split_i8_s<2>(t, g, s); break;
      case 3: launch_split_i8_s<3>(t, g, s); break;
      case 4: launch_split_i8_s<4>(t, g, s); break;
      case 6: launch_split_i8_s<6>(t, g, s); break;
      default: launch_split_i8_s<5>(t, g, s); break;
    }
    SPLIT_HIP(hipGetLastError());
    return TNT_OK;
  }
  SplitGemmArgs g{t.jobs, t.tiles, d, ap, bp, ae, be, alpha, beta};
  switch (S) {
    case 2: launch_digits_s<2>(pa, items, s); launch_split_s<2>(t, g, s); break;
    case 3: launch_digits_s<3>(pa, items, s); launch_split_s<3>(t, g, s); break;
    case 4: launch_digits_s<4>(pa, items, s); launch_split_s<4>(t, g, s); break;
    case 6: launch_digits_s<6>(pa, items, s); launch_split_s<6>(t, g, s); break;
    default: launch_digits_s<5>(pa, items, s); launch_split_s<5>(t, g, s); break;
  }
  SPLIT_HIP(hipGetLastError());
  return TNT_OK;
}

}  // namespace tnt_dev
