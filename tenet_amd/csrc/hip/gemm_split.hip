// Split-BF16 (Ozaki-style) block GEMM for gfx950.
//
// FP64 C = op(A) op(B) from BF16 matrix-core products:
//   1. per row i of op(A) an exponent e_i with max_k |A_ik| < 2^e_i, per
//      column j of op(B) an exponent f_j (pre-pass);
//   2. a = A_ik 2^-e_i in (-1, 1) is cut into S base-256 digits by repeated
//      round-to-nearest: d_1 in [-256, 256], d_p in [-128, 128] for p > 1 --
//      small integers, exact in BF16 (pre-pass, written as S BF16 planes,
//      k-contiguous per row / column, zero-padded to a multiple of 32);
//   3. C_ij = 2^(e_i+f_j) sum_{p+q<=S+1} 2^-8(p+q) sum_k dA_p,ik dB_q,kj.
//      Products of one level l = p+q go to one FP32 accumulator.  Within a
//      K chunk of 128 every partial sum is an integer of magnitude
//      < 1.75 * 2^16 * 128 < 2^24, so the FP32 MFMA accumulation is exact;
//      after each chunk the level sums are scaled by 2^-8l and added in FP64.
// The only rounding is the truncation of the digit expansion (relative
// 2^-8S per operand) and the dropped levels l > S+1; S = 5 gives relative
// Frobenius errors around 1e-12 on random data (tests hold it to 1e-10).
//
// MFMA: v_mfma_f32_16x16x32_bf16.  Both operands use the same k-to-lane
// assignment (k = 8*(lane/16) + 0..7 within a 32-wide step), so the result
// does not depend on the hardware's internal k order; rows/cols map to
// lane % 16 and D follows the 16x16 layout (row 4*(lane/16)+r, col lane%16).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>

#include "../../../include/tenet_hip.h"
#include "split.hpp"

namespace tnt_dev {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SPLIT_KS = 32;     // k per LDS stage (one MFMA k-step)
constexpr int SPLIT_CHUNK = 128; // exact FP32 accumulation window
constexpr int SPLIT_BM = 64, SPLIT_BN = 64;
constexpr int SPLIT_LDS_PAD = 8;  // bf16 elements of row padding

struct SplitJob {
  int64_t d_off;     // element offset of C from dbase
  int64_t a_row0;    // plane line base: row i of this job starts at (a_row0 + i) * kp
  int64_t b_col0;    // plane line base of the columns of op(B)
  int64_t a_exp0;    // first row of this job in the A exponent array
  int64_t b_exp0;    // first column of this job in the B exponent array
  int32_t m, n, kp;  // kp = k rounded up to SPLIT_KS
  int32_t ldd;
};

struct SplitTables {
  SplitJob* jobs = nullptr;
  GTile* tiles = nullptr;
  int64_t* prefix = nullptr;   // per-job start of its workgroups in the flat pre-pass grids
  uint32_t n_tiles = 0;
  uint32_t exp_blocks_a = 0, exp_blocks_b = 0, slice_blocks_a = 0, slice_blocks_b = 0;
  int64_t a_elems = 0, b_elems = 0, a_rows = 0, b_cols = 0;
};

struct SplitWorkspace {
  void* planes = nullptr;  // A planes | B planes (bf16) | A exps | B exps (int32)
  size_t bytes = 0;
  std::unordered_map<std::string, SplitTables> tables;
  std::vector<void*> table_allocs;
  ~SplitWorkspace() {
    if (planes) (void)hipFree(planes);
    for (auto p : table_allocs) (void)hipFree(p);
  }
};

SplitWorkspace* split_workspace_create() { return new SplitWorkspace(); }
void split_workspace_destroy(SplitWorkspace* ws) { delete ws; }

// ---------------------------------------------------------------------------
// Pre-passes.  Both run over a flat grid; a workgroup finds its job by a
// binary search of the per-job prefix of workgroup counts.
//   element (line, kk): side 0 -> op(A)[line][kk], side 1 -> op(B)[kk][line]
// ---------------------------------------------------------------------------
struct PrepArgs {
  const GJob* jobs;
  const SplitJob* sj;
  const int64_t* prefix;  // n_jobs + 1 entries
  int n_jobs;
  uintptr_t base;         // operand base (A or B)
  int32_t* exps;          // exponents of this side
  __bf16* planes;         // planes of this side
  int64_t plane_stride;
  int side;
  int slices;
};

__device__ __forceinline__ int find_job(const int64_t* prefix, int n, int64_t b) {
  int lo = 0, hi = n;  // prefix[lo] <= b < prefix[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (prefix[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool line_contiguous(const GJob& g, int side) {
  // identity op(A) (column-major rows) and adjoint op(B) are contiguous along lines
  return side == 0 ? !(g.ops & 1) : ((g.ops & 2) != 0);
}

__device__ __forceinline__ int64_t elem_index(const GJob& g, int side, int line, int kk) {
  if (side == 0) return (g.ops & 1) ? (kk + (int64_t)line * g.lda) : (line + (int64_t)kk * g.lda);
  return (g.ops & 2) ? (line + (int64_t)kk * g.ldb) : (kk + (int64_t)line * g.ldb);
}

// Exponent per line: 256 lines per workgroup (line-contiguous sources, one
// thread per line) or 4 lines per workgroup (k-contiguous, one wave per line).
__global__ __launch_bounds__(256) void split_exponents(PrepArgs x) {
  const int j = find_job(x.prefix, x.n_jobs, blockIdx.x);
  const GJob g = x.jobs[j];
  const SplitJob s = x.sj[j];
  const int64_t local = blockIdx.x - x.prefix[j];
  const int lines = x.side == 0 ? g.m : g.n;
  const int64_t exp0 = x.side == 0 ? s.a_exp0 : s.b_exp0;
  const double* M = reinterpret_cast<const double*>(x.base) + (x.side == 0 ? g.a_off : g.b_off);
  double mx = 0.0;
  int line;
  if (line_contiguous(g, x.side)) {
    line = (int)(local * 256 + threadIdx.x);
    if (line >= lines) return;
    for (int kk = 0; kk < g.k; ++kk) mx = fmax(mx, fabs(M[elem_index(g, x.side, line, kk)]));
  } else {
    const int lane = threadIdx.x & 63;
    line = (int)(local * 4 + (threadIdx.x >> 6));
    if (line >= lines) return;
    for (int kk = lane; kk < g.k; kk += 64) mx = fmax(mx, fabs(M[elem_index(g, x.side, line, kk)]));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
    if (lane != 0) return;
  }
  int e = 0;
  if (mx > 0.0) frexp(mx, &e);  // mx = f * 2^e, f in [0.5, 1): |x| < 2^e
  x.exps[exp0 + line] = e;
}

// Digit planes: one 64x64 (line x k) tile per workgroup, staged through LDS
// so that both the FP64 read and the BF16 writes are coalesced.
__global__ __launch_bounds__(256) void split_slices(PrepArgs x) {
  __shared__ double T[64][65];
  const int j = find_job(x.prefix, x.n_jobs, blockIdx.x);
  const GJob g = x.jobs[j];
  const SplitJob s = x.sj[j];
  const int64_t local = blockIdx.x - x.prefix[j];
  const int lines = x.side == 0 ? g.m : g.n;
  const int ktiles = (s.kp + 63) / 64;
  const int l0 = (int)(local / ktiles) * 64, k0 = (int)(local % ktiles) * 64;
  const int64_t line0 = x.side == 0 ? s.a_row0 : s.b_col0;
  const int64_t exp0 = x.side == 0 ? s.a_exp0 : s.b_exp0;
  const double* M = reinterpret_cast<const double*>(x.base) + (x.side == 0 ? g.a_off : g.b_off);
  const int t = threadIdx.x;
  const bool lc = line_contiguous(g, x.side);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    int li, ki;
    if (lc) { li = t & 63; ki = (t >> 6) + 4 * r; }
    else { ki = t & 63; li = (t >> 6) + 4 * r; }
    const int line = l0 + li, kk = k0 + ki;
    double v = 0.0;
    if (line < lines && kk < g.k) v = ldexp(M[elem_index(g, x.side, line, kk)], -x.exps[exp0 + line]);
    T[li][ki] = v;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int ki = t & 63, li = (t >> 6) + 4 * r;
    const int line = l0 + li, kk = k0 + ki;
    if (line >= lines || kk >= s.kp) continue;
    double v = T[li][ki];
    __bf16* out = x.planes + (line0 + line) * (int64_t)s.kp + kk;
    for (int p = 0; p < x.slices; ++p) {
      v *= 256.0;
      const double d = rint(v);
      v -= d;
      out[p * x.plane_stride] = (__bf16)(float)d;
    }
  }
}

// ---------------------------------------------------------------------------
// Split GEMM kernel: 64x64 tile per workgroup, 4 waves (2x2) of 32x32.
// ---------------------------------------------------------------------------
struct SplitGemmArgs {
  const SplitJob* sj;
  const GTile* tiles;
  uintptr_t dbase;
  const __bf16* a_planes;
  const __bf16* b_planes;
  const int32_t* a_exp;
  const int32_t* b_exp;
  int64_t a_stride, b_stride;  // plane strides (elements)
  double alpha, beta;
};

template <int S>
__global__ __launch_bounds__(256) void split_gemm_kernel(SplitGemmArgs g) {
  constexpr int LDK = SPLIT_KS + SPLIT_LDS_PAD;
  __shared__ __attribute__((aligned(16))) __bf16 As[S][SPLIT_BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[S][SPLIT_BN][LDK];
  const GTile tile = g.tiles[blockIdx.x];
  const SplitJob job = g.sj[tile.job];
  const int m0 = tile.tm * SPLIT_BM, n0 = tile.tn * SPLIT_BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  const int r16 = lane & 15, kq = (lane >> 4) * 8;

  f32x4 acc[S][2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc64[i][j][r] = 0.0;

  const int nsteps = job.kp / SPLIT_KS;
  const int steps_per_chunk = SPLIT_CHUNK / SPLIT_KS;
  for (int step = 0; step < nsteps; ++step) {
    if (step % steps_per_chunk == 0) {
#pragma unroll
      for (int l = 0; l < S; ++l)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) acc[l][i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    // stage S planes of a 64x32 A slab and a 64x32 B slab: 16-byte vectors,
    // 4 per row, 64 rows, S planes -> S*256 vectors per operand
    const int k0 = step * SPLIT_KS;
    __syncthreads();
#pragma unroll
    for (int p = 0; p < S; ++p) {
      const int row = tid >> 2, seg = (tid & 3) * 8;
      const int gi = m0 + row, gj = n0 + row;
      uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
      if (gi < job.m)
        va = *reinterpret_cast<const uint4*>(g.a_planes + p * g.a_stride + (job.a_row0 + gi) * (int64_t)job.kp + k0 + seg);
      if (gj < job.n)
        vb = *reinterpret_cast<const uint4*>(g.b_planes + p * g.b_stride + (job.b_col0 + gj) * (int64_t)job.kp + k0 + seg);
      *reinterpret_cast<uint4*>(&As[p][row][seg]) = va;
      *reinterpret_cast<uint4*>(&Bs[p][row][seg]) = vb;
    }
    __syncthreads();
    bf16x8 bf[S][2];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < 2; ++j) bf[q][j] = *reinterpret_cast<const bf16x8*>(&Bs[q][wn + j * 16 + r16][kq]);
#pragma unroll
    for (int p = 0; p < S; ++p) {
      bf16x8 af[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) af[i] = *reinterpret_cast<const bf16x8*>(&As[p][wm + i * 16 + r16][kq]);
#pragma unroll
      for (int q = 0; q < S - p; ++q) {  // level (p+1)+(q+1) <= S+1
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[q][j], acc[p + q][i][j], 0, 0, 0);
      }
    }
    if ((step + 1) % steps_per_chunk == 0 || step + 1 == nsteps) {
#pragma unroll
      for (int l = 0; l < S; ++l) {
        const double sc = ldexp(1.0, -8 * (l + 2));
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc64[i][j][r] = fma((double)acc[l][i][j][r], sc, acc64[i][j][r]);
      }
    }
  }
  // epilogue
  double* D = reinterpret_cast<double*>(g.dbase) + job.d_off;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int col = n0 + wn + j * 16 + r16;
    if (col >= job.n) continue;
    const int fcol = g.b_exp[job.b_exp0 + col];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (row < job.m) {
          const double v = g.alpha * ldexp(acc64[i][j][r], g.a_exp[job.a_exp0 + row] + fcol);
          double* p = D + row + (int64_t)col * job.ldd;
          *p = (g.beta != 0.0) ? v + g.beta * *p : v;
        }
      }
    }
  }
}

static int split_slices_setting = 5;

int split_set_slices(int slices) {
  if (slices < 2 || slices > 6) return TNT_E_INVALID;
  split_slices_setting = slices;
  return TNT_OK;
}

static void launch_split_gemm(int S, const SplitGemmArgs& a, uint32_t n, hipStream_t s) {
  switch (S) {
    case 2: hipLaunchKernelGGL(split_gemm_kernel<2>, dim3(n), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(split_gemm_kernel<3>, dim3(n), dim3(256), 0, s, a); break;
    case 4: hipLaunchKernelGGL(split_gemm_kernel<4>, dim3(n), dim3(256), 0, s, a); break;
    case 6: hipLaunchKernelGGL(split_gemm_kernel<6>, dim3(n), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL(split_gemm_kernel<5>, dim3(n), dim3(256), 0, s, a); break;
  }
}

#define SPLIT_HIP(expr)                                                    \
  do {                                                                     \
    hipError_t _e = (expr);                                                \
    if (_e != hipSuccess) {                                                \
      err = std::string(#expr) + ": " + hipGetErrorString(_e);             \
      return _e == hipErrorOutOfMemory ? TNT_E_OOM : TNT_E_HIP;            \
    }                                                                      \
  } while (0)

int split_gemm_run(SplitWorkspace* ws, const std::vector<GJob>& host_jobs, const GJob* dev_jobs,
                   const std::vector<GTile>& /*host_tiles*/, const GTile* /*dev_tiles*/, uintptr_t d,
                   uintptr_t a, uintptr_t b, double alpha, double beta, hipStream_t s, bool capturing,
                   std::string& err) {
  const int S = split_slices_setting;
  if (host_jobs.empty()) return TNT_OK;
  std::string key(reinterpret_cast<const char*>(host_jobs.data()), host_jobs.size() * sizeof(GJob));
  key.push_back((char)S);
  auto it = ws->tables.find(key);
  if (it == ws->tables.end()) {
    if (capturing) {
      err = "split-BF16 tables are built on first use; run the call once before capturing";
      return TNT_E_INVALID;
    }
    SplitTables t;
    std::vector<SplitJob> sj(host_jobs.size());
    std::vector<GTile> tiles;
    std::vector<int64_t> pre_ea{0}, pre_eb{0}, pre_sa{0}, pre_sb{0};
    for (size_t i = 0; i < host_jobs.size(); ++i) {
      const GJob& j = host_jobs[i];
      SplitJob x{};
      x.d_off = j.d_off;
      x.m = j.m;
      x.n = j.n;
      x.kp = ((std::max(j.k, 1) + SPLIT_KS - 1) / SPLIT_KS) * SPLIT_KS;
      x.ldd = j.ldd;
      // planes are [line][kp] per job, concatenated: row0 counts elements / kp
      // is not uniform across jobs, so store element offsets and rows separately
      x.a_row0 = t.a_elems;  // element offset (converted below)
      x.b_col0 = t.b_elems;
      t.a_elems += (int64_t)j.m * x.kp;
      t.b_elems += (int64_t)j.n * x.kp;
      sj[i] = x;
      bool a_lc = !(j.ops & 1), b_lc = (j.ops & 2) != 0;
      pre_ea.push_back(pre_ea.back() + (a_lc ? (j.m + 255) / 256 : (j.m + 3) / 4));
      pre_eb.push_back(pre_eb.back() + (b_lc ? (j.n + 255) / 256 : (j.n + 3) / 4));
      int kt = (x.kp + 63) / 64;
      pre_sa.push_back(pre_sa.back() + (int64_t)((j.m + 63) / 64) * kt);
      pre_sb.push_back(pre_sb.back() + (int64_t)((j.n + 63) / 64) * kt);
      for (int tm = 0; tm < (j.m + SPLIT_BM - 1) / SPLIT_BM; ++tm)
        for (int tn = 0; tn < (j.n + SPLIT_BN - 1) / SPLIT_BN; ++tn) tiles.push_back({(int32_t)i, tm, tn, 0});
    }
    // Exponents and planes are indexed by a per-side "line" counter; planes
    // address line * kp, so give every job a line base with base * kp equal
    // to its element offset: round the element offset up to a multiple of kp.
    int64_t ae = 0, be = 0, ar = 0, bc = 0;
    std::vector<int64_t> exp_row(sj.size()), exp_col(sj.size());
    for (size_t i = 0; i < sj.size(); ++i) {
      SplitJob& x = sj[i];
      ae = (ae + x.kp - 1) / x.kp * x.kp;
      be = (be + x.kp - 1) / x.kp * x.kp;
      x.a_row0 = ae / x.kp;
      x.b_col0 = be / x.kp;
      ae += (int64_t)x.m * x.kp;
      be += (int64_t)x.n * x.kp;
      exp_row[i] = ar;
      exp_col[i] = bc;
      x.a_exp0 = ar;
      x.b_exp0 = bc;
      ar += x.m;
      bc += x.n;
    }
    t.a_elems = ae;
    t.b_elems = be;
    t.a_rows = ar;
    t.b_cols = bc;
    (void)exp_row;
    (void)exp_col;
    t.n_tiles = (uint32_t)tiles.size();
    t.exp_blocks_a = (uint32_t)pre_ea.back();
    t.exp_blocks_b = (uint32_t)pre_eb.back();
    t.slice_blocks_a = (uint32_t)pre_sa.back();
    t.slice_blocks_b = (uint32_t)pre_sb.back();
    std::vector<int64_t> pre;
    pre.insert(pre.end(), pre_ea.begin(), pre_ea.end());
    pre.insert(pre.end(), pre_eb.begin(), pre_eb.end());
    pre.insert(pre.end(), pre_sa.begin(), pre_sa.end());
    pre.insert(pre.end(), pre_sb.begin(), pre_sb.end());
    void *pj, *pt, *pp;
    SPLIT_HIP(hipMalloc(&pj, sizeof(SplitJob) * sj.size()));
    ws->table_allocs.push_back(pj);
    SPLIT_HIP(hipMalloc(&pt, sizeof(GTile) * std::max<size_t>(tiles.size(), 1)));
    ws->table_allocs.push_back(pt);
    SPLIT_HIP(hipMalloc(&pp, sizeof(int64_t) * pre.size()));
    ws->table_allocs.push_back(pp);
    SPLIT_HIP(hipMemcpyAsync(pj, sj.data(), sizeof(SplitJob) * sj.size(), hipMemcpyHostToDevice, s));
    if (!tiles.empty()) SPLIT_HIP(hipMemcpyAsync(pt, tiles.data(), sizeof(GTile) * tiles.size(), hipMemcpyHostToDevice, s));
    SPLIT_HIP(hipMemcpyAsync(pp, pre.data(), sizeof(int64_t) * pre.size(), hipMemcpyHostToDevice, s));
    SPLIT_HIP(hipStreamSynchronize(s));
    t.jobs = static_cast<SplitJob*>(pj);
    t.tiles = static_cast<GTile*>(pt);
    t.prefix = static_cast<int64_t*>(pp);
    it = ws->tables.emplace(key, t).first;
  }
  const SplitTables& t = it->second;
  // workspace: S A planes + S B planes (bf16), exponents per plane line (int32)
  const size_t bytes = 2 * (size_t)S * (size_t)(t.a_elems + t.b_elems) + 4 * (size_t)(t.a_rows + t.b_cols) + 256;
  if (bytes > ws->bytes) {
    if (capturing) {
      err = "split-BF16 workspace must be sized before graph capture; run the call once first";
      return TNT_E_INVALID;
    }
    SPLIT_HIP(hipStreamSynchronize(s));
    if (ws->planes) SPLIT_HIP(hipFree(ws->planes));
    ws->planes = nullptr;
    ws->bytes = 0;
    SPLIT_HIP(hipMalloc(&ws->planes, bytes));
    ws->bytes = bytes;
  }
  __bf16* ap = static_cast<__bf16*>(ws->planes);
  __bf16* bp = ap + (size_t)S * t.a_elems;
  int32_t* ae = reinterpret_cast<int32_t*>(bp + (size_t)S * t.b_elems);
  int32_t* be = ae + t.a_rows;
  const int n_jobs = (int)host_jobs.size();
  const int64_t* pre = t.prefix;
  PrepArgs pa{dev_jobs, t.jobs, pre, n_jobs, a, ae, ap, t.a_elems, 0, S};
  PrepArgs pb{dev_jobs, t.jobs, pre + (n_jobs + 1), n_jobs, b, be, bp, t.b_elems, 1, S};
  if (t.exp_blocks_a) hipLaunchKernelGGL(split_exponents, dim3(t.exp_blocks_a), dim3(256), 0, s, pa);
  if (t.exp_blocks_b) hipLaunchKernelGGL(split_exponents, dim3(t.exp_blocks_b), dim3(256), 0, s, pb);
  pa.prefix = pre + 2 * (n_jobs + 1);
  pb.prefix = pre + 3 * (n_jobs + 1);
  if (t.slice_blocks_a) hipLaunchKernelGGL(split_slices, dim3(t.slice_blocks_a), dim3(256), 0, s, pa);
  if (t.slice_blocks_b) hipLaunchKernelGGL(split_slices, dim3(t.slice_blocks_b), dim3(256), 0, s, pb);
  SplitGemmArgs g{t.jobs, t.tiles, d, ap, bp, ae, be, t.a_elems, t.b_elems, alpha, beta};
  if (t.n_tiles) launch_split_gemm(S, g, t.n_tiles, s);
  SPLIT_HIP(hipGetLastError());
  return TNT_OK;
}

}  // namespace tnt_dev
