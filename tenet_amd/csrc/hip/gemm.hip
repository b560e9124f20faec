// Grouped block GEMM for gfx950 on the matrix cores.
//
// One launch covers every (job, tile) of a batch of uneven coupled-sector
// GEMMs (the reference issues one GEMM per sector:
// tenet-operations/src/fusion_replay.rs:627-734, tenet-dense/src/executor.rs:217-312).
// Column-major operands; op = identity or adjoint (transpose) per operand.
//
//   FP64: v_mfma_f64_16x16x4f64, FP64 accumulate.
//   FP32: operands rounded to FP32 while staging into LDS,
//         v_mfma_f32_16x16x4f32, FP32 accumulate, FP64 epilogue.
// Payload storage S is FP64 or FP32 (DenseRead::F32 jobs of the reference's
// executor seam); loads convert to the arithmetic type, the epilogue computes
// alpha acc + beta D in FP64 and rounds to S once.
//
// Workgroup = 4 wave64s in a 2x2 grid (8 in a 2x4 grid for TILE_128x128_W8);
// each wave owns a (BM/GM)x(BN/GN) sub-tile made of 16x16 MFMA tiles.  K advances in steps of
// BK=16 through LDS with a register prefetch of the next K slab, so global
// loads of slab k+1 overlap the MFMAs of slab k.
//
// MFMA 16x16x4 lane mapping (one value per lane for A and B, four for D),
// established on gfx950 with a one-hot probe of each builtin:
//   A[i][k]: i = lane % 16, k = lane / 16
//   B[k][j]: j = lane % 16, k = lane / 16
//   D[i][j]: j = lane % 16, and for register r = 0..3
//            FP64 (v_mfma_f64_16x16x4f64):  i = lane / 16 + 4 * r
//            FP32 (v_mfma_f32_16x16x4f32):  i = 4 * (lane / 16) + r
#include "kernels.hpp"

#include <algorithm>
#include <type_traits>

namespace tnt_dev {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaOp;

template <>
struct MfmaOp<double> {
  using acc_t = f64x4;
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

template <>
struct MfmaOp<float> {
  using acc_t = f32x4;
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int BK = 16;

// Workgroup = GM x GN waves (wave w at row w % GM, column w / GM of the tile).
// Register budget: 4-wave tiles at most 256 VGPRs+AGPRs per lane, so two
// waves share each SIMD (one wave per SIMD leaves LDS and global latency
// exposed); 8-wave tiles and 64 x 64 tiles at most 128, four waves per SIMD
// (for 64 x 64 tiles four workgroups per CU, as many as their LDS allows).
//
// Every global load of the main loop is unconditional, so the loop has no
// branch but its trip test and the compiler keeps each K slab's loads in
// flight across the slab's MFMAs (predicated edge loads had made it wait for
// the A slab right after issuing it):
//  * a tile that would cross the bottom (right) edge of a job with m >= BM
//    (n >= BN) is moved back inside the job and stores only the rows
//    (columns) from its own start on, which no other tile stores; the rows it
//    recomputes are the padding an edge tile computes anyway;
//  * in a job smaller than the tile, line indices are clamped to the last
//    line (clamped lines only feed outputs that are never stored);
//  * the K remainder is the FIRST slab, loaded in the prologue with the
//    elements past it zeroed, so every slab of the loop is a full BK.
// The adjoint flags are template parameters (a launch covers tiles of one
// op pair), and the LDS fragments of k-step ks+1 are read while the MFMAs of
// k-step ks issue.
template <typename T, int BM, int BN, int GM = 2, int GN = 2, bool BETA = false, typename S = double,
          bool AADJ = false, bool BADJ = false>
__global__ __launch_bounds__(64 * GM * GN) __attribute__((amdgpu_waves_per_eu(GM * GN == 8 || (BM == 64 && BN == 64) ? 4 : 2))) void gemm_kernel(
    GemmArgs g) {
  constexpr int NTH = 64 * GM * GN;
  constexpr int WM = BM / GM, WN = BN / GN;
  constexpr int MT = WM / 16, NT = WN / 16;
  constexpr int PA = 2, PB = 2;  // LDS row padding (elements)
  constexpr int A_LOADS = BM * BK / NTH, B_LOADS = BN * BK / NTH;
  static_assert(A_LOADS * NTH == BM * BK && B_LOADS * NTH == BN * BK, "tile loads must split evenly");
  __shared__ T As[2][BK][BM + PA];
  __shared__ T Bs[2][BK][BN + PB];

  const GTile tile = g.tiles[blockIdx.x];
  const GJob job = g.jobs[tile.job];
  const int m = job.m, n = job.n, k = job.k;
  const int m0o = tile.tm * BM, n0o = tile.tn * BN;  // first row / column this tile stores
  const int m0 = m >= BM ? min(m0o, m - BM) : 0;     // tile origin, inside the job when it fits
  const int n0 = n >= BN ? min(n0o, n - BN) : 0;
  const gptr<const S> A = gmem<const S>(g.abase) + job.a_off;
  const gptr<const S> B = gmem<const S>(g.bbase) + job.b_off;
  const gptr<S> D = gmem<S>(g.dbase) + job.d_off;
  const int64_t lda = job.lda, ldb = job.ldb, ldd = job.ldd;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = (wave % GM) * WM, wn = (wave / GM) * WN;

  using acc_t = typename MfmaOp<T>::acc_t;
  acc_t acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};

  // ---- loader geometry (computed once) ----
  // A identity: contiguous along i (C = BM, O = BK); adjoint: along k (C = BK, O = BM)
  // B identity: contiguous along k (C = BK, O = BN); adjoint: along j (C = BN, O = BK)
  // Load r of a thread reads element (line, kk) of the slab: A identity
  // (i = a_c, kk = a_o + a_step r), adjoint (i = a_o + a_step r, kk = a_c);
  // B identity (j = b_o + b_step r, kk = b_c), adjoint (j = b_c, kk = b_o + b_step r).
  constexpr int a_step = AADJ ? NTH / BK : NTH / BM;
  constexpr int b_step = BADJ ? NTH / BN : NTH / BK;
  const int a_c = AADJ ? tid % BK : tid % BM;
  const int a_o = AADJ ? tid / BK : tid / BM;
  const int b_c = BADJ ? tid % BN : tid % BK;
  const int b_o = BADJ ? tid / BN : tid / BK;
  // element pointers at the load's own kk of slab 0: one per load where the
  // line changes with r (adjoint A, identity B), else one per thread
  constexpr int NPA = AADJ ? A_LOADS : 1, NPB = BADJ ? 1 : B_LOADS;
  gptr<const S> pa[NPA], pb[NPB];
  if constexpr (AADJ) {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) pa[r] = A + (int64_t)min(m0 + a_o + a_step * r, m - 1) * lda + a_c;
  } else {
    pa[0] = A + min(m0 + a_c, m - 1) + (int64_t)a_o * lda;
  }
  if constexpr (BADJ) {
    pb[0] = B + min(n0 + b_c, n - 1) + (int64_t)b_o * ldb;
  } else {
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) pb[r] = B + (int64_t)min(n0 + b_o + b_step * r, n - 1) * ldb + b_c;
  }
  auto aptr = [&](int r) { return AADJ ? pa[r] : pa[0] + (int64_t)r * a_step * lda; };
  auto bptr = [&](int r) { return BADJ ? pb[0] + (int64_t)r * b_step * ldb : pb[r]; };
  auto akk = [&](int r) { return AADJ ? a_c : a_o + a_step * r; };
  auto bkk = [&](int r) { return BADJ ? b_o + b_step * r : b_c; };
  // element offset of k index kk from k index 0 of the same line
  auto aoff = [&](int kk) { return AADJ ? (int64_t)kk : (int64_t)kk * lda; };
  auto boff = [&](int kk) { return BADJ ? (int64_t)kk * ldb : (int64_t)kk; };

  // Slab registers: one set, or two for the two-slab prefetch of 64 x 64
  // tiles (their slab is 1024 MFMA cycles per wave, four workgroups per CU:
  // one slab ahead left global latency exposed; 85 against 81 % of peak on
  // 4096^3, 65 against 62 % on the U(1) environment's A^dagger.T2 jobs).
  constexpr bool PF2 = BM == 64 && BN == 64;
  S ra[A_LOADS], rb[B_LOADS], ra1[PF2 ? A_LOADS : 1], rb1[PF2 ? B_LOADS : 1];
  auto store = [&](int buf, const S (&xa)[A_LOADS], const S (&xb)[B_LOADS]) {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) {
      if (!AADJ) As[buf][akk(r)][a_c] = (T)xa[r];
      else As[buf][a_c][a_o + a_step * r] = (T)xa[r];
    }
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) {
      if (!BADJ) Bs[buf][b_c][b_o + b_step * r] = (T)xb[r];
      else Bs[buf][bkk(r)][b_c] = (T)xb[r];
    }
  };
  // the slab at k offsets (ak_, bk_) from each load's own line start
  auto load = [&](S (&xa)[A_LOADS], S (&xb)[B_LOADS], int64_t ak_, int64_t bk_) {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) xa[r] = aptr(r)[ak_];
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) xb[r] = bptr(r)[bk_];
  };
  T fa[2][MT], fb[2][NT];
  auto frags = [&](int buf, int ks, int s) {
    const int kk = ks * 4 + (lane >> 4);
#pragma unroll
    for (int i = 0; i < MT; ++i) fa[s][i] = As[buf][kk][wm + i * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < NT; ++j) fb[s][j] = Bs[buf][kk][wn + j * 16 + (lane & 15)];
  };
  auto mfmas = [&](int s) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) acc[i][j] = MfmaOp<T>::run(fa[s][i], fb[s][j], acc[i][j]);
  };

  // Prologue: slab 0 holds k in [0, kf), kf = the K remainder (1..BK); its
  // elements at kk >= kf are zero (their addresses clamped to kf - 1).
  // Nothing contracted (k == 0): no slab, no load; the epilogue gives beta D.
  const int kf = k > 0 ? k - (k - 1) / BK * BK : 0;
  const int nk = k > 0 ? (k - kf) / BK + 1 : 0;
  if (nk > 0) {
#pragma unroll
    for (int r = 0; r < A_LOADS; ++r) {
      const int kk = akk(r);
      const S v = aptr(r)[aoff(min(kk, kf - 1) - kk)];
      ra[r] = kk < kf ? v : S(0);
    }
#pragma unroll
    for (int r = 0; r < B_LOADS; ++r) {
      const int kk = bkk(r);
      const S v = bptr(r)[boff(min(kk, kf - 1) - kk)];
      rb[r] = kk < kf ? v : S(0);
    }
    store(0, ra, rb);
  }
  // Slab t >= 1 covers k in [kf + (t - 1) BK, kf + t BK).
  int64_t ak = aoff(kf), bk = boff(kf);
  const int64_t a_kadv = aoff(BK), b_kadv = boff(BK);
  if constexpr (PF2) {
    if (nk > 1) load(ra1, rb1, ak, bk);  // slab 1
    ak += a_kadv;
    bk += b_kadv;
  }
  __syncthreads();
  // A wave whose part of the tile stores nothing (wholly outside the job, or
  // inside the rows / columns a moved edge tile only recomputes) skips its
  // LDS reads and MFMAs -- the waves of the other workgroups on its SIMD get
  // the matrix core -- but keeps loading its share of the slabs and meets
  // every barrier.
  const bool live = __builtin_amdgcn_readfirstlane((m0 + wm + WM > m0o) && (m0 + wm < m) &&
                                                   (n0 + wn + WN > n0o) && (n0 + wn < n));
  if (live && nk > 0) frags(0, 0, 0);
  auto compute = [&](int cur) {
    if (live) {
#pragma unroll
      for (int ks = 0; ks < BK / 4; ++ks) {
        if (ks + 1 < BK / 4) frags(cur, ks + 1, (ks + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
        mfmas(ks & 1);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  };
  if constexpr (!PF2) {
    // Double-buffered LDS: slab t is read from buffer t&1 while slab t+1 is
    // fetched into registers and then written to the other buffer; one
    // barrier per slab.
    for (int t = 0; t < nk; ++t) {
      const int cur = t & 1;
      if (t + 1 < nk) {
        load(ra, rb, ak, bk);
        ak += a_kadv;
        bk += b_kadv;
      }
      compute(cur);
      if (t + 1 < nk) store(cur ^ 1, ra, rb);
      __syncthreads();
      if (live && t + 1 < nk) frags(cur ^ 1, 0, 0);
    }
  } else if (nk == 1) {
    compute(0);  // the remainder slab only: no further load, no LDS write
  } else if (nk > 1) {
    // Two slabs in flight: while slab t is multiplied from LDS buffer t&1,
    // slab t+1 waits in one register set (loaded an iteration earlier) and
    // slab t+2 is fetched into the other; after the MFMAs slab t+1 goes to
    // the other LDS buffer.  Past the last slab the loads re-read slab
    // nk - 1 (unconditional; the values are not used).
    const int64_t a_last = aoff(kf) + (int64_t)(nk - 2) * a_kadv, b_last = boff(kf) + (int64_t)(nk - 2) * b_kadv;
    auto step = [&](int t, int cur, const S (&nxa)[A_LOADS], const S (&nxb)[B_LOADS], S (&isa)[A_LOADS],
                    S (&isb)[B_LOADS]) {
      const bool more = t + 2 < nk;
      load(isa, isb, more ? ak : a_last, more ? bk : b_last);
      ak += a_kadv;
      bk += b_kadv;
      compute(cur);
      if (t + 1 < nk) store(cur ^ 1, nxa, nxb);
      __syncthreads();
      if (live && t + 1 < nk) frags(cur ^ 1, 0, 0);
    };
    for (int t = 0; t < nk; t += 2) {
      step(t, 0, ra1, rb1, ra, rb);
      if (t + 1 >= nk) break;
      step(t + 1, 1, ra, rb, ra1, rb1);
    }
  }
  if (!live) return;
  const double alpha = g.alpha, beta = g.beta;
  if (g.cmap) {
    // Output transform folded in: core column col -> dst[off + row * rs],
    // scaled by c, in the same operation order as the recoupling pass
    // ((alpha c) x + beta y), so the result is that of transform-after-GEMM.
    const ColMapEntry* __restrict__ cm = g.cmap + g.cmap_start[tile.job];
    const gptr<S> D0 = gmem<S>(g.dbase);
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      if (col < n0o || col >= n) continue;
      const ColMapEntry ce = cm[col];
      if (ce.off < 0) continue;
      const double ac = alpha * ce.c;
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm + i * 16 + MfmaOp<T>::row(lane, r);
          if (row >= m0o && row < m) {
            const gptr<S> p = D0 + ce.off + (int64_t)row * ce.rs;
            if constexpr (BETA) {
              const double w = (double)*p;
              *p = (S)(ac * (double)acc[i][j][r] + beta * w);
            } else {
              *p = (S)(ac * (double)acc[i][j][r]);
            }
          }
        }
    }
    return;
  }

  // Epilogue: D = alpha * acc (+ beta * D in the BETA instance, launched for
  // beta != 0) over rows [m0o, m) and columns [n0o, n) of the tile.  There
  // the D values of one row of fragments are read unconditionally (indices
  // clamped into the job) before any is stored, so the wave waits once per
  // fragment row rather than once per element; the beta == 0 instance keeps
  // its registers for the main loop.
#pragma unroll
  for (int i = 0; i < MT; ++i) {
    double old[BETA ? NT : 1][4];
    if constexpr (BETA) {
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int col = min(n0 + wn + j * 16 + (lane & 15), n - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          old[j][r] = (double)D[min(m0 + wm + i * 16 + MfmaOp<T>::row(lane, r), m - 1) + (int64_t)col * ldd];
      }
    }
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int col = n0 + wn + j * 16 + (lane & 15);
      if (col < n0o || col >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm + i * 16 + MfmaOp<T>::row(lane, r);
        if (row >= m0o && row < m) {
          const double v = alpha * (double)acc[i][j][r];
          if constexpr (BETA) D[row + (int64_t)col * ldd] = (S)(v + beta * old[j][r]);
          else D[row + (int64_t)col * ldd] = (S)v;
        }
      }
    }
  }
}

// Narrow-output GEMM (n <= 16): memory-bound, so no matrix cores -- one
// thread per output row, op(B) staged in LDS, FP64 FMA.  Reads op(A) once
// (coalesced along rows for identity A) and writes D once.
template <typename T, typename S = double>
__global__ __launch_bounds__(256) void gemm_skinny_kernel(GemmArgs g) {
  constexpr int NMAX = 16, KC = 256;
  __shared__ double Bsm[KC][NMAX];
  const GTile tile = g.tiles[blockIdx.x];
  const GJob job = g.jobs[tile.job];
  const int m = job.m, n = job.n, k = job.k;
  const bool a_adj = job.ops & 1, b_adj = job.ops & 2;
  const gptr<const S> A = gmem<const S>(g.abase) + job.a_off;
  const gptr<const S> B = gmem<const S>(g.bbase) + job.b_off;
  const gptr<S> D = gmem<S>(g.dbase) + job.d_off;
  const int i = tile.tm * 256 + threadIdx.x;
  double acc[NMAX];
#pragma unroll
  for (int j = 0; j < NMAX; ++j) acc[j] = 0.0;
  for (int k0 = 0; k0 < k; k0 += KC) {
    const int kc = min(KC, k - k0);
    __syncthreads();
    for (int e = threadIdx.x; e < kc * NMAX; e += 256) {
      const int kk = e / NMAX, j = e % NMAX;
      double v = 0.0;
      if (j < n) v = (double)(b_adj ? B[j + (int64_t)(k0 + kk) * job.ldb] : B[(k0 + kk) + (int64_t)j * job.ldb]);
      Bsm[kk][j] = (double)(T)v;
    }
    __syncthreads();
    if (i < m) {
      for (int kk = 0; kk < kc; ++kk) {
        const int gk = k0 + kk;
        const double a = (double)(T)(a_adj ? A[gk + (int64_t)i * job.lda] : A[i + (int64_t)gk * job.lda]);
#pragma unroll
        for (int j = 0; j < NMAX; ++j) acc[j] = fma(a, Bsm[kk][j], acc[j]);
      }
    }
  }
  if (i < m) {
    // For beta != 0 the row of D is read (columns clamped) before any store.
    const double alpha = g.alpha, beta = g.beta;
    double old[NMAX] = {};
    if (beta != 0.0) {
#pragma unroll
      for (int j = 0; j < NMAX; ++j) old[j] = (double)D[i + (int64_t)min(j, n - 1) * job.ldd];
    }
#pragma unroll
    for (int j = 0; j < NMAX; ++j) {
      if (j < n) {
        const double v = alpha * acc[j];
        D[i + (int64_t)j * job.ldd] = (S)((beta != 0.0) ? v + beta * old[j] : v);
      }
    }
  }
}

// Small-job GEMM: each wave of the workgroup computes its own 32 x 32 tile
// (2 x 2 MFMA fragments) with operands loaded straight from global memory into
// registers -- no LDS, no barriers, waves independent.  For the many tiny
// sector GEMMs of batched contractions (median 30 x 16 x 30), where a 64 x 64
// workgroup tile is mostly padding and the launch is bound by memory latency.
// Inside a block of 8 k the lanes of group q = lane / 16 supply k = k0 + q + 4p
// to the p-th MFMA.  Loads are unconditional (indices clamped into the job's
// operands); out-of-range k is zeroed by select, so the loop body is
// straight-line code and each stage keeps its own registers.  A job with
// nothing contracted (k == 0) loads nothing: its operands may be empty.
template <typename T, bool BETA = false, typename S = double>
__global__ __launch_bounds__(256) void gemm_small_kernel(GemmArgs g, uint32_t n_tiles) {
  constexpr int MT = 2, NT = 2;
  const int lane = threadIdx.x & 63, l16 = lane & 15, q4 = lane >> 4;
  const uint32_t ti = blockIdx.x * 4u + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ti >= n_tiles) return;
  const GTile tile = g.tiles[ti];
  const GJob job = g.jobs[tile.job];
  const int m = job.m, n = job.n, k = job.k;
  const int m0 = tile.tm * 32, n0 = tile.tn * 32;
  const bool a_adj = job.ops & 1, b_adj = job.ops & 2;
  const int64_t sAi = a_adj ? job.lda : 1, sAk = a_adj ? 1 : job.lda;
  const int64_t sBj = b_adj ? 1 : job.ldb, sBk = b_adj ? job.ldb : 1;
  const gptr<const S> A = gmem<const S>(g.abase) + job.a_off;
  const gptr<const S> B = gmem<const S>(g.bbase) + job.b_off;
  gptr<const S> pa[MT], pb[NT];
#pragma unroll
  for (int i = 0; i < MT; ++i) pa[i] = A + (int64_t)min(m0 + 16 * i + l16, m - 1) * sAi;
#pragma unroll
  for (int j = 0; j < NT; ++j) pb[j] = B + (int64_t)min(n0 + 16 * j + l16, n - 1) * sBj;
  using acc_t = typename MfmaOp<T>::acc_t;
  acc_t acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = acc_t{0, 0, 0, 0};
  S ra[2][2][MT], rb[2][2][NT];  // [stage][p][fragment]
  const int nb = (k + 7) / 8;
  auto load = [&](auto sc, int blk) {
    constexpr int s = decltype(sc)::value;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int kk = min(blk * 8 + q4 + 4 * p, k - 1);
#pragma unroll
      for (int i = 0; i < MT; ++i) ra[s][p][i] = pa[i][kk * sAk];
#pragma unroll
      for (int j = 0; j < NT; ++j) rb[s][p][j] = pb[j][kk * sBk];
    }
  };
  auto compute = [&](auto sc, int blk) {
    constexpr int s = decltype(sc)::value;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const bool ok = blk * 8 + q4 + 4 * p < k;
      T af[MT], bf[NT];
#pragma unroll
      for (int i = 0; i < MT; ++i) af[i] = (T)(ok ? ra[s][p][i] : S(0));
#pragma unroll
      for (int j = 0; j < NT; ++j) bf[j] = (T)(ok ? rb[s][p][j] : S(0));
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = MfmaOp<T>::run(af[i], bf[j], acc[i][j]);
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  // two register stages: block t + 1 is in flight while block t computes; the
  // scheduling barriers keep the compiler from hoisting the next stage's loads
  // above the current compute (which would make it wait for them)
  if (k > 0) load(S0{}, 0);
  for (int t = 0; t < nb; t += 2) {
    load(S1{}, t + 1);
    __builtin_amdgcn_sched_barrier(0);
    compute(S0{}, t);
    __builtin_amdgcn_sched_barrier(0);
    load(S0{}, t + 2);
    __builtin_amdgcn_sched_barrier(0);
    compute(S1{}, t + 1);
    __builtin_amdgcn_sched_barrier(0);
  }
  // Epilogue.  The BETA instance (launched for beta != 0) reads the whole
  // tile of D first, indices clamped into the job, so the wave waits once
  // before the stores instead of once per element.
  const gptr<S> D = gmem<S>(g.dbase) + job.d_off;
  const double alpha = g.alpha, beta = g.beta;
  double old[BETA ? MT : 1][NT][4];
  if constexpr (BETA) {
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        const int col = min(n0 + 16 * j + l16, n - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r)
          old[i][j][r] = (double)D[min(m0 + 16 * i + MfmaOp<T>::row(lane, r), m - 1) + (int64_t)col * job.ldd];
      }
  }
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int col = n0 + 16 * j + l16;
      if (col >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 16 * i + MfmaOp<T>::row(lane, r);
        if (row < m) {
          const double v = alpha * (double)acc[i][j][r];
          if constexpr (BETA) D[row + (int64_t)col * job.ldd] = (S)(v + beta * old[i][j][r]);
          else D[row + (int64_t)col * job.ldd] = (S)v;
        }
      }
    }
}

// The gathered small GEMM of a tiny contraction: gemm_small_kernel's tile
// walk and MFMA sequence (so a job's sums are formed in the same order), with
// an operand element e of a transformed operand read as map[e].c *
// source[map[e].idx] -- the value the Single-block replay would have written
// to the transformed operand -- and a result element stored through the output
// map, alpha * c * acc (+ beta * dst), as the output replay would.  Workgroups
// past the tile workgroups set the fix-up elements to beta * dst.  Indices are
// clamped into each job's operands as in gemm_small_kernel, and every map
// index was checked against its payload at upload.
template <bool LM, bool RM, bool OM, bool BETA>
__global__ __launch_bounds__(256) void gemm_tiny_kernel(GemmArgs g, TinyArgs t, uint32_t n_tiles) {
  if (blockIdx.x >= t.tile_blocks) {
    const uint32_t nb = gridDim.x - t.tile_blocks;
    for (uint32_t i = (blockIdx.x - t.tile_blocks) * 256u + threadIdx.x; i < t.n_fix; i += nb * 256u) {
      double* p = t.dst + t.fix[i];
      *p = BETA ? g.beta * *p : 0.0;
    }
    return;
  }
  constexpr int MT = 2, NT = 2;
  const int lane = threadIdx.x & 63, l16 = lane & 15, q4 = lane >> 4;
  const uint32_t ti = blockIdx.x * 4u + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (ti >= n_tiles) return;
  const GTile tile = g.tiles[ti];
  const GJob job = g.jobs[tile.job];
  const int m = job.m, n = job.n, k = job.k;
  const int m0 = tile.tm * 32, n0 = tile.tn * 32;
  const bool a_adj = job.ops & 1, b_adj = job.ops & 2;
  const int64_t sAi = a_adj ? job.lda : 1, sAk = a_adj ? 1 : job.lda;
  const int64_t sBj = b_adj ? 1 : job.ldb, sBk = b_adj ? job.ldb : 1;
  const gptr<const double> A = gmem<const double>(g.abase);
  const gptr<const double> B = gmem<const double>(g.bbase);
  int64_t ea[MT], eb[NT];  // element offsets of the fragment rows / columns
#pragma unroll
  for (int i = 0; i < MT; ++i) ea[i] = job.a_off + (int64_t)min(m0 + 16 * i + l16, m - 1) * sAi;
#pragma unroll
  for (int j = 0; j < NT; ++j) eb[j] = job.b_off + (int64_t)min(n0 + 16 * j + l16, n - 1) * sBj;
  auto lval = [&](int64_t e) -> double {
    if constexpr (LM) {
      const TinyEntry te = t.lmap[e];
      return te.c * A[te.idx];
    } else {
      return A[e];
    }
  };
  auto rval = [&](int64_t e) -> double {
    if constexpr (RM) {
      const TinyEntry te = t.rmap[e];
      return te.c * B[te.idx];
    } else {
      return B[e];
    }
  };
  f64x4 acc[MT][NT];
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) acc[i][j] = f64x4{0, 0, 0, 0};
  double ra[2][2][MT], rb[2][2][NT];  // [stage][p][fragment]
  const int nb = (k + 7) / 8;
  auto load = [&](auto sc, int blk) {
    constexpr int s = decltype(sc)::value;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int kk = min(blk * 8 + q4 + 4 * p, k - 1);
#pragma unroll
      for (int i = 0; i < MT; ++i) ra[s][p][i] = lval(ea[i] + kk * sAk);
#pragma unroll
      for (int j = 0; j < NT; ++j) rb[s][p][j] = rval(eb[j] + kk * sBk);
    }
  };
  auto compute = [&](auto sc, int blk) {
    constexpr int s = decltype(sc)::value;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const bool ok = blk * 8 + q4 + 4 * p < k;
      double af[MT], bf[NT];
#pragma unroll
      for (int i = 0; i < MT; ++i) af[i] = ok ? ra[s][p][i] : 0.0;
#pragma unroll
      for (int j = 0; j < NT; ++j) bf[j] = ok ? rb[s][p][j] : 0.0;
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[i][j] = MfmaOp<double>::run(af[i], bf[j], acc[i][j]);
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  if (k > 0 && k <= 8 * kTinyKBlocks) {
    // short K: every operand fragment of the job in flight at once (one
    // memory round trip instead of one per 16 k); loads past K re-read the
    // last k (clamped) and their products are skipped, so the sums are the
    // pipelined loop's
    double fa[kTinyKBlocks][2][MT], fb[kTinyKBlocks][2][NT];
#pragma unroll
    for (int blk = 0; blk < kTinyKBlocks; ++blk)
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int kk = min(blk * 8 + q4 + 4 * p, k - 1);
#pragma unroll
        for (int i = 0; i < MT; ++i) fa[blk][p][i] = lval(ea[i] + kk * sAk);
#pragma unroll
        for (int j = 0; j < NT; ++j) fb[blk][p][j] = rval(eb[j] + kk * sBk);
      }
#pragma unroll
    for (int blk = 0; blk < kTinyKBlocks; ++blk) {
      if (blk >= nb) break;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const bool ok = blk * 8 + q4 + 4 * p < k;
        double af[MT], bf[NT];
#pragma unroll
        for (int i = 0; i < MT; ++i) af[i] = ok ? fa[blk][p][i] : 0.0;
#pragma unroll
        for (int j = 0; j < NT; ++j) bf[j] = ok ? fb[blk][p][j] : 0.0;
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
          for (int j = 0; j < NT; ++j) acc[i][j] = MfmaOp<double>::run(af[i], bf[j], acc[i][j]);
      }
    }
  } else {
    if (k > 0) load(S0{}, 0);
    for (int tt = 0; tt < nb; tt += 2) {
      load(S1{}, tt + 1);
      __builtin_amdgcn_sched_barrier(0);
      compute(S0{}, tt);
      __builtin_amdgcn_sched_barrier(0);
      load(S0{}, tt + 2);
      __builtin_amdgcn_sched_barrier(0);
      compute(S1{}, tt + 1);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  const double alpha = g.alpha, beta = g.beta;
  const gptr<double> D = gmem<double>(g.dbase);
#pragma unroll
  for (int i = 0; i < MT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j) {
      const int col = n0 + 16 * j + l16;
      if (col >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + 16 * i + MfmaOp<double>::row(lane, r);
        if (row >= m) continue;
        const int64_t e = job.d_off + row + (int64_t)col * job.ldd;
        if constexpr (OM) {
          const TinyEntry te = t.omap[e];
          if (te.idx >= 0) {
            double* p = t.dst + te.idx;
            const double c = alpha * te.c;
            const double v = acc[i][j][r];
            *p = BETA ? c * v + beta * *p : c * v;
          }
        } else {
          const double v = alpha * acc[i][j][r];
          D[e] = BETA ? v + beta * D[e] : v;
        }
      }
    }
}

template <bool LM, bool RM, bool OM>
static void launch_tiny3(const GemmArgs& g, const TinyArgs& t, uint32_t grid, uint32_t n_tiles, hipStream_t s) {
  if (g.beta != 0.0)
    hipLaunchKernelGGL((gemm_tiny_kernel<LM, RM, OM, true>), dim3(grid), dim3(256), 0, s, g, t, n_tiles);
  else
    hipLaunchKernelGGL((gemm_tiny_kernel<LM, RM, OM, false>), dim3(grid), dim3(256), 0, s, g, t, n_tiles);
}

void launch_gemm_tiny(const GemmArgs& g, const TinyArgs& t0, uint32_t n_tiles, hipStream_t s) {
  TinyArgs t = t0;
  t.tile_blocks = (n_tiles + 3) / 4;
  const uint32_t fix_blocks = t.n_fix ? std::min<uint32_t>(64u, (t.n_fix + 255) / 256) : 0u;
  const uint32_t grid = t.tile_blocks + fix_blocks;
  if (grid == 0) return;
  const int key = (t.lmap ? 1 : 0) | (t.rmap ? 2 : 0) | (t.omap ? 4 : 0);
  switch (key) {
    case 0: launch_tiny3<false, false, false>(g, t, grid, n_tiles, s); break;
    case 1: launch_tiny3<true, false, false>(g, t, grid, n_tiles, s); break;
    case 2: launch_tiny3<false, true, false>(g, t, grid, n_tiles, s); break;
    case 3: launch_tiny3<true, true, false>(g, t, grid, n_tiles, s); break;
    case 4: launch_tiny3<false, false, true>(g, t, grid, n_tiles, s); break;
    case 5: launch_tiny3<true, false, true>(g, t, grid, n_tiles, s); break;
    case 6: launch_tiny3<false, true, true>(g, t, grid, n_tiles, s); break;
    default: launch_tiny3<true, true, true>(g, t, grid, n_tiles, s); break;
  }
}

template <typename T, bool BETA, typename S, bool AADJ, bool BADJ>
static void launch_tiles(const GemmArgs& a, int cfg, uint32_t n_tiles, hipStream_t s) {
  switch (cfg) {
    case TILE_128x128:
      hipLaunchKernelGGL((gemm_kernel<T, 128, 128, 2, 2, BETA, S, AADJ, BADJ>), dim3(n_tiles), dim3(256), 0, s, a);
      break;
    case TILE_128x64:
      hipLaunchKernelGGL((gemm_kernel<T, 128, 64, 2, 2, BETA, S, AADJ, BADJ>), dim3(n_tiles), dim3(256), 0, s, a);
      break;
    case TILE_64x128:
      hipLaunchKernelGGL((gemm_kernel<T, 64, 128, 2, 2, BETA, S, AADJ, BADJ>), dim3(n_tiles), dim3(256), 0, s, a);
      break;
    case TILE_128x128_W8:
      hipLaunchKernelGGL((gemm_kernel<T, 128, 128, 2, 4, BETA, S, AADJ, BADJ>), dim3(n_tiles), dim3(512), 0, s, a);
      break;
    default:
      hipLaunchKernelGGL((gemm_kernel<T, 64, 64, 2, 2, BETA, S, AADJ, BADJ>), dim3(n_tiles), dim3(256), 0, s, a);
      break;
  }
}

template <typename T, bool BETA, typename S>
static void launch_tiles_ops(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  switch (ops & 3) {
    case 0: launch_tiles<T, BETA, S, false, false>(a, cfg, n_tiles, s); break;
    case 1: launch_tiles<T, BETA, S, true, false>(a, cfg, n_tiles, s); break;
    case 2: launch_tiles<T, BETA, S, false, true>(a, cfg, n_tiles, s); break;
    default: launch_tiles<T, BETA, S, true, true>(a, cfg, n_tiles, s); break;
  }
}

// ops: the op pair (GJob::ops) shared by every tile of the launch; the
// narrow-output and one-wave kernels read each job's own ops.
template <typename T, typename S>
static void launch_cfg(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  if (n_tiles == 0) return;
  switch (cfg) {
    case TILE_SKINNY:
      hipLaunchKernelGGL((gemm_skinny_kernel<T, S>), dim3(n_tiles), dim3(256), 0, s, a);
      break;
    case TILE_SMALL:  // four independent one-wave tiles per workgroup
      if (a.beta != 0.0)
        hipLaunchKernelGGL((gemm_small_kernel<T, true, S>), dim3((n_tiles + 3) / 4), dim3(256), 0, s, a, n_tiles);
      else
        hipLaunchKernelGGL((gemm_small_kernel<T, false, S>), dim3((n_tiles + 3) / 4), dim3(256), 0, s, a, n_tiles);
      break;
    default:
      if (a.beta != 0.0) launch_tiles_ops<T, true, S>(a, cfg, ops, n_tiles, s);
      else launch_tiles_ops<T, false, S>(a, cfg, ops, n_tiles, s);
      break;
  }
}

void launch_gemm_f64(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  launch_cfg<double, double>(a, cfg, ops, n_tiles, s);
}

void launch_gemm_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  launch_cfg<float, double>(a, cfg, ops, n_tiles, s);
}

void launch_gemm_f64_on_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  launch_cfg<double, float>(a, cfg, ops, n_tiles, s);
}

void launch_gemm_f32_on_f32(const GemmArgs& a, int cfg, int ops, uint32_t n_tiles, hipStream_t s) {
  launch_cfg<float, float>(a, cfg, ops, n_tiles, s);
}

}  // namespace tnt_dev
