// Recoupling replay kernel for gfx950: Single (scaled strided copy), Multi
// (per-element small dense recoupling  dst_j = sum_i U[j,i] src_i, i.e. the
// reference's pack -> src * U^T -> scatter fused into one pass with FP64
// coefficients) and Scale (inactive destination blocks).
//
// Semantics follow tenet-operations/src/transform_replay.rs:3413-3606 (Single
// :4428-4505, pack :4697-4743, recoupling GEMM :4519-4565, scatter :4745-4822,
// inactive scaling :288-330).  The kernel never materialises the packed
// scratch: every source element is read once and every destination element
// written once, so a launch moves exactly the algorithmic bytes.
//
// Work decomposition: each normalised block is a set of "planes" (dims 0 and
// 1, destination order) indexed by an outer index over dims 2..rank-1.  A
// work item covers [e_begin, e_begin + e_count) of the plane for outer
// indices [o_begin, o_begin + o_count).  The outer index is wave-uniform, so
// per-layout base offsets live in scalar registers; lanes sweep the plane
// destination-fastest (coalesced writes) with an incremental (x, y) walk.
// Single items of one outer index are also flattened at upload into XFlat
// records (plane base offsets and strides), read with one scalar load per
// item; the Single-class kernel runs them before the Single items of several
// outer indices, in the same launch.
#include "kernels.hpp"

#include <algorithm>
#include <cstdlib>

namespace tnt_dev {

static_assert(sizeof(XItem) == 24, "XItem layout");

constexpr int kUnroll = 4;
static_assert(kSingleItem == 256 * kSingleUnroll, "Single items are swept 256 elements at a time");

__device__ __forceinline__ int64_t base_offset(const XLayout& L, const int64_t* idx, int rank) {
  int64_t off = L.offset;
#pragma unroll
  for (int d = 2; d < kMaxRank; ++d)
    if (d < rank) off += idx[d] * L.strides[d];
  return off;
}

__device__ __forceinline__ void outer_index(const XBlock& b, uint64_t o, int64_t* idx) {
  const int rank = b.rank;
#pragma unroll
  for (int d = 2; d < kMaxRank; ++d) {
    if (d < rank) {
      idx[d] = (int64_t)(o % (uint64_t)b.dims[d]);
      o /= (uint64_t)b.dims[d];
    } else {
      idx[d] = 0;
    }
  }
}

// Lanes of the workgroup: waves split outer indices when there are enough of
// them, else the plane.
struct ItemSweep {
  uint64_t o_first, o_step, o_end;
  uint32_t e_lo, e_hi;
  __device__ ItemSweep(const XItem& it, int wave) {
    if (it.o_count >= 4) {
      o_first = it.o_begin + wave;
      o_step = 4;
      e_lo = it.e_begin;
      e_hi = it.e_begin + it.e_count;
    } else {
      o_first = it.o_begin;
      o_step = 1;
      const uint32_t q = (it.e_count + 3) / 4;
      e_lo = it.e_begin + min(it.e_count, q * wave);
      e_hi = it.e_begin + min(it.e_count, q * (wave + 1));
    }
    o_end = it.o_begin + it.o_count;
  }
};

// Single (scaled strided copy) and Scale (inactive destination) items.
__device__ void single_item(const XformArgs& a, const XItem it) {
  const XBlock& b = a.blocks[it.block];
  const int rank = b.rank;
  const int64_t n0 = b.dims[0];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  if (it.o_count == 1 && b.kind == XK_SINGLE) {
    // one outer index, up to kSingleItem plane elements: element
    // e = e_begin + 64 * (4u + wave) + lane, u < kSingleUnroll -- the whole
    // workgroup sweeps the item front to back in 256-element steps, with all
    // kSingleUnroll loads of a lane in flight before its stores
    int64_t idx[kMaxRank];
    outer_index(b, it.o_begin, idx);
    const XLayout& Ld = a.layouts[b.lay];
    const XLayout& Ls = a.layouts[b.lay + 1];
    const int64_t od = base_offset(Ld, idx, rank), os = base_offset(Ls, idx, rank);
    const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
    const int64_t s0 = Ls.strides[0], s1 = rank > 1 ? Ls.strides[1] : 0;
    const double c = a.alpha * a.coeffs[b.coeff];
    const bool readd = a.overwrite == 0 && a.beta != 0.0;
    const uint32_t n0u = (uint32_t)n0;
    // (x, y) of the lane's first element, then steps of 256 elements
    const uint32_t f0 = 64u * wave + lane;
    uint32_t x = (it.e_begin + f0) % n0u, y = (it.e_begin + f0) / n0u;
    const uint32_t sx = 256u % n0u, sy = 256u / n0u;
    double v[kSingleUnroll];
    int64_t dp[kSingleUnroll];
    bool ok[kSingleUnroll];
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u) {
      ok[u] = f0 + 256u * u < it.e_count;  // unconditional loads (see flat_item)
      v[u] = a.src[ok[u] ? os + x * s0 + y * s1 : os];
      dp[u] = od + x * d0 + y * d1;
      x += sx;
      y += sy;
      if (x >= n0u) {
        x -= n0u;
        y += 1;
      }
    }
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u) {
      if (ok[u]) {
        double* p = a.dst + dp[u];
        *p = readd ? c * v[u] + a.beta * *p : c * v[u];
      }
    }
    return;
  }

  const ItemSweep w(it, wave);
  // incremental plane walk: e -> (x, y) with e = x + n0 * y; step 64
  const int64_t dx = 64 % n0, dy = 64 / n0;
  const double alpha = a.alpha, beta = a.beta;
  const bool readd = a.overwrite == 0 && beta != 0.0;

  for (uint64_t o = w.o_first; o < w.o_end; o += w.o_step) {
    int64_t idx[kMaxRank];
    outer_index(b, o, idx);
    uint32_t e = w.e_lo + lane;
    int64_t x = e % n0, y = e / n0;
    if (b.kind == XK_SINGLE) {
      const XLayout& Ld = a.layouts[b.lay];
      const XLayout& Ls = a.layouts[b.lay + 1];
      const int64_t od = base_offset(Ld, idx, rank), os = base_offset(Ls, idx, rank);
      const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
      const int64_t s0 = Ls.strides[0], s1 = rank > 1 ? Ls.strides[1] : 0;
      const double c = alpha * a.coeffs[b.coeff];
      const double* __restrict__ src = a.src;
      double* __restrict__ dst = a.dst;
      // four independent loads in flight per lane before their stores; the
      // loads are unconditional (past the end a lane re-reads element (0, 0))
      while (e < w.e_hi) {
        double v[kUnroll];
        int64_t dp[kUnroll];
        bool ok[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          ok[u] = e < w.e_hi;
          v[u] = src[ok[u] ? os + x * s0 + y * s1 : os];
          dp[u] = od + x * d0 + y * d1;
          e += 64;
          x += dx;
          y += dy;
          if (x >= n0) { x -= n0; y += 1; }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (ok[u]) {
            double* p = dst + dp[u];
            *p = readd ? c * v[u] + beta * *p : c * v[u];
          }
        }
      }
    } else {  // XK_SCALE: inactive destination block
      const XLayout& Ld = a.layouts[b.lay];
      const int64_t od = base_offset(Ld, idx, rank);
      const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
      for (; e < w.e_hi; e += 64) {
        double* p = a.dst + od + x * d0 + y * d1;
        *p = readd ? beta * *p : 0.0;
        x += dx;
        y += dy;
        if (x >= n0) { x -= n0; y += 1; }
      }
    }
  }
}

// One flattened Single item: element e = e_begin + 64 * (4u + wave) + lane,
// u < kSingleUnroll.  The loads are unconditional (a lane past the item reads
// plane element (0, 0) and discards it), so all of a lane's loads are in
// flight before its first store; a predicated load would be waited for at
// the branch merge.
__device__ __forceinline__ void flat_item(const XformArgs& a, const XFlat& f, double alpha, double beta,
                                          bool readd) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const double c = alpha * f.c;
  const uint32_t f0 = 64u * wave + lane;
  uint32_t x = (f.e_begin + f0) % f.n0, y = (f.e_begin + f0) / f.n0;
  const uint32_t sx = 256u % f.n0, sy = 256u / f.n0;
  double v[kSingleUnroll];
  int64_t dp[kSingleUnroll];
  bool ok[kSingleUnroll];
#pragma unroll
  for (int u = 0; u < kSingleUnroll; ++u) {
    ok[u] = f0 + 256u * u < f.e_count;
    const int64_t sp = ok[u] ? f.os + x * f.s0 + y * f.s1 : f.os;
    dp[u] = ok[u] ? f.od + x * f.d0 + y * f.d1 : f.od;
    v[u] = a.src[sp];
    x += sx;
    y += sy;
    if (x >= f.n0) {
      x -= f.n0;
      y += 1;
    }
  }
  if (readd) {
    double w[kSingleUnroll];
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u) w[u] = a.dst[dp[u]];
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u)
      if (ok[u]) a.dst[dp[u]] = c * v[u] + beta * w[u];
  } else {
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u)
      if (ok[u]) a.dst[dp[u]] = c * v[u];
  }
}

// Multi items with at most NS source layouts; dst_j = sum_i U[j, i] src_i is
// accumulated with fma in source order.  Few sources (NS <= 4, e.g. the fused
// narrow contractions): the source layouts of the current outer index are
// held in registers and each lane recouples UN plane elements (e, e + 64,
// ..., e + 64 (UN - 1)) per step, so UN * nsrc loads are in flight before the
// stores.  More sources: one element per lane per step, layouts read from the
// (scalar-cached) table, which keeps the register budget -- and the
// occupancy -- of the wide groups.
template <int NS, int UN = 2>
__device__ void multi_item(const XformArgs& a, const XItem it, const double* Ublock = nullptr) {
  constexpr bool kFew = NS <= 4;
  static_assert(UN >= 1 && UN <= 8, "elements per lane and step");
  const XBlock& b = a.blocks[it.block];
  const int rank = b.rank;
  const int64_t n0 = b.dims[0];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nd = b.ndst, ns = b.nsrc;
  const XLayout* __restrict__ L = a.layouts + b.lay;
  const double* __restrict__ U = Ublock ? Ublock : a.coeffs + b.coeff;
  const double* __restrict__ src = a.src;
  double* __restrict__ dst = a.dst;
  const double alpha = a.alpha, beta = a.beta;
  const bool readd = a.overwrite == 0 && beta != 0.0;
  const ItemSweep w(it, wave);
  const int64_t dx = 64 % n0, dy = 64 / n0;

  for (uint64_t o = w.o_first; o < w.o_end; o += w.o_step) {
    int64_t idx[kMaxRank];
    outer_index(b, o, idx);
    uint32_t e = w.e_lo + lane;
    int64_t x = e % n0, y = e / n0;
    if constexpr (kFew) {
      int64_t sb[NS], sx[NS], sy[NS];
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        if (i < ns) {
          const XLayout& Ls = L[nd + i];
          sb[i] = base_offset(Ls, idx, rank);
          sx[i] = Ls.strides[0];
          sy[i] = rank > 1 ? Ls.strides[1] : 0;
        } else {
          sb[i] = sx[i] = sy[i] = 0;
        }
      }
      while (e < w.e_hi) {
        // plane coordinates of elements e + 64 u; past the end of the range a
        // load repeats element e (unconditional loads) and nothing is stored
        int64_t xs[UN], ys[UN];
        bool ok[UN];
        xs[0] = x;
        ys[0] = y;
        ok[0] = true;
#pragma unroll
        for (int u = 1; u < UN; ++u) {
          xs[u] = xs[u - 1] + dx;
          ys[u] = ys[u - 1] + dy;
          if (xs[u] >= n0) { xs[u] -= n0; ys[u] += 1; }
          ok[u] = e + 64u * u < w.e_hi;
        }
        double v[UN][NS];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
          for (int i = 0; i < NS; ++i) {
            v[u][i] = 0.0;
            if (i < ns) v[u][i] = src[ok[u] ? sb[i] + xs[u] * sx[i] + ys[u] * sy[i] : sb[i] + x * sx[i] + y * sy[i]];
          }
        for (int j = 0; j < nd; ++j) {
          const XLayout& Ld = L[j];
          const int64_t db = base_offset(Ld, idx, rank);
          const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
          double acc[UN];
#pragma unroll
          for (int u = 0; u < UN; ++u) acc[u] = 0.0;
#pragma unroll
          for (int i = 0; i < NS; ++i) {
            if (i < ns) {
              const double c = U[j * ns + i];
#pragma unroll
              for (int u = 0; u < UN; ++u) acc[u] = fma(c, v[u][i], acc[u]);
            }
          }
#pragma unroll
          for (int u = 0; u < UN; ++u) {
            if (ok[u]) {
              double* p = dst + db + xs[u] * d0 + ys[u] * d1;
              *p = readd ? alpha * acc[u] + beta * *p : alpha * acc[u];
            }
          }
        }
        e += 64u * UN;
        x = xs[UN - 1] + dx;
        y = ys[UN - 1] + dy;
        if (x >= n0) { x -= n0; y += 1; }
      }
    } else {
      for (; e < w.e_hi; e += 64) {
        double v[NS];
#pragma unroll
        for (int i = 0; i < NS; ++i) {
          if (i < ns) {
            const XLayout& Ls = L[nd + i];
            v[i] = src[base_offset(Ls, idx, rank) + x * Ls.strides[0] + (rank > 1 ? y * Ls.strides[1] : 0)];
          } else {
            v[i] = 0.0;
          }
        }
        for (int j = 0; j < nd; ++j) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < NS; ++i)
            if (i < ns) acc = fma(U[j * ns + i], v[i], acc);
          const XLayout& Ld = L[j];
          double* p = dst + base_offset(Ld, idx, rank) + x * Ld.strides[0] + (rank > 1 ? y * Ld.strides[1] : 0);
          *p = readd ? alpha * acc + beta * *p : alpha * acc;
        }
        x += dx;
        y += dy;
        if (x >= n0) { x -= n0; y += 1; }
      }
    }
  }
}

// Transposed Single tile: read a 64 x 64 tile along the source-contiguous
// axis (dims[1]) into LDS, write it along the destination-contiguous axis
// (dims[0]); both global passes are coalesced.  Row pitch 65 keeps the
// column reads of the write pass free of bank conflicts.
__device__ void transpose_item(const XformArgs& a, const XItem it, double (&tile)[kTransposeTile][kTransposeTile + 1]) {
  constexpr int T = kTransposeTile;
  const XBlock& b = a.blocks[it.block];
  const int rank = b.rank;
  const int64_t n0 = b.dims[0], n1 = b.dims[1];
  const int64_t ntx = (n0 + T - 1) / T;
  const int64_t x0 = (int64_t)(it.e_begin % ntx) * T, y0 = (int64_t)(it.e_begin / ntx) * T;
  int64_t idx[kMaxRank];
  outer_index(b, it.o_begin, idx);
  const XLayout& Ld = a.layouts[b.lay];
  const XLayout& Ls = a.layouts[b.lay + 1];
  const int64_t od = base_offset(Ld, idx, rank), os = base_offset(Ls, idx, rank);
  const int64_t d0 = Ld.strides[0], d1 = Ld.strides[1];
  const int64_t s0 = Ls.strides[0], s1 = Ls.strides[1];
  const double c = a.alpha * a.coeffs[b.coeff];
  const bool readd = a.overwrite == 0 && a.beta != 0.0;
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  {
    // all T / 4 loads of a lane in flight before the LDS writes
    const int64_t y = y0 + l;
    double v[T / 4];
#pragma unroll
    for (int r = 0; r < T / 4; ++r) {
      const int64_t x = x0 + w + 4 * r;
      v[r] = (x < n0 && y < n1) ? a.src[os + x * s0 + y * s1] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < T / 4; ++r) tile[w + 4 * r][l] = v[r];
  }
  __syncthreads();
  {
    const int64_t x = x0 + l;
#pragma unroll
    for (int r = 0; r < T / 4; ++r) {
      const int64_t y = y0 + w + 4 * r;
      if (x < n0 && y < n1) {
        double* p = a.dst + od + x * d0 + y * d1;
        const double v = c * tile[l][w + 4 * r];
        *p = readd ? v + a.beta * *p : v;
      }
    }
  }
  __syncthreads();  // the tile is refilled by the next item
}

// One kernel per item class, so each class gets the register budget (and
// occupancy) of its own path.  Persistent grid-stride loop over the class's
// items: concurrently resident workgroups process consecutive items, so the
// waves of the whole GPU sweep one contiguous window of the payloads at a
// time (DRAM locality), like a plain grid-stride copy.
__global__ __launch_bounds__(256) void xform_single_kernel(XformArgs a, uint32_t begin, uint32_t n) {
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) single_item(a, a.items[begin + i]);
}

// The whole Single class in one launch: the n_flat flattened items (one
// record each, no block / layout lookups) first, then the n Single / Scale
// items of several outer indices from [begin, begin + n) of the item table.
__global__ __launch_bounds__(256) void xform_single_all_kernel(XformArgs a, uint32_t begin, uint32_t n,
                                                               uint32_t n_flat) {
  const double alpha = a.alpha, beta = a.beta;
  const bool readd = a.overwrite == 0 && beta != 0.0;
  const uint32_t total = n_flat + n;
  for (uint32_t i = blockIdx.x; i < total; i += gridDim.x) {
    if (i < n_flat) flat_item(a, a.flat[i], alpha, beta, readd);
    else single_item(a, a.items[begin + (i - n_flat)]);
  }
}

__global__ __launch_bounds__(256) void xform_transpose_kernel(XformArgs a, uint32_t begin, uint32_t n) {
  __shared__ double tile[kTransposeTile][kTransposeTile + 1];
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) transpose_item(a, a.items[begin + i], tile);
}

template <int NS, int UN = 2>
__global__ __launch_bounds__(256) void xform_multi_kernel(XformArgs a, uint32_t begin, uint32_t n) {
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) multi_item<NS, UN>(a, a.items[begin + i]);
}

// Few-source Multi items with their coefficients gathered per item: the
// workgroup sums the block's ndst x nsrc coefficients into LDS (each as
// coeff_gather_kernel forms it), then recouples the item from there.
__global__ __launch_bounds__(256) void xform_multi_gather_kernel(XformArgs a, uint32_t begin, uint32_t n) {
  __shared__ double su[kMaxGatherCoef];
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const XItem it = a.items[begin + i];
    const XBlock& b = a.blocks[it.block];
    const int nc = b.ndst * b.nsrc;
    for (int t = threadIdx.x; t < nc; t += 256) {
      const int64_t c = (int64_t)b.coeff + t;
      const int64_t p0 = a.gptr[c], p1 = a.gptr[c + 1];
      double u = p0 < p1 ? a.gscale[p0] * a.gsrc[a.gidx[p0]] : 0.0;  // one term: the plain product
      for (int64_t p = p0 + 1; p < p1; ++p) u = fma(a.gscale[p], a.gsrc[a.gidx[p]], u);
      su[t] = u;
    }
    __syncthreads();
    multi_item<4, 4>(a, it, su);
    __syncthreads();  // su is refilled by the next item
  }
}

void launch_xform_multi_gather(const XformArgs& a, uint32_t begin, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  const uint32_t grid = n < 4096u ? n : 4096u;
  hipLaunchKernelGGL(xform_multi_gather_kernel, dim3(grid), dim3(256), 0, s, a, begin, n);
}

// Plane elements per lane and step of the few-source Multi launches
// (TENET_AMD_MULTI_UNROLL = 2 or 4, read once; for A/B runs).
static int multi_unroll() {
  static const int u = [] {
    const char* e = std::getenv("TENET_AMD_MULTI_UNROLL");
    return e && std::atoi(e) == 2 ? 2 : 4;
  }();
  return u;
}

// Every class in one launch (small plans: one launch costs more than the
// occupancy the per-class kernels buy).  Items of one workgroup are uniform,
// so the transpose barrier is reached by the whole workgroup.
__global__ __launch_bounds__(256) void xform_any_kernel(XformArgs a, uint32_t n) {
  __shared__ double tile[kTransposeTile][kTransposeTile + 1];
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const XItem it = a.items[i];
    const XBlock& b = a.blocks[it.block];
    switch (xclass_of(b.kind, b.nsrc)) {
      case XC_SINGLE:
        single_item(a, it);
        break;
      case XC_TRANSPOSE:
        transpose_item(a, it, tile);
        break;
      case XC_MULTI4:
        multi_item<4>(a, it);
        break;
      case XC_MULTI16:
        multi_item<16>(a, it);
        break;
      default:
        multi_item<kMaxMultiSrc>(a, it);
        break;
    }
  }
}

__global__ __launch_bounds__(256) void coeff_gather_kernel(double* coeffs, const double* src, const int64_t* ptr,
                                                          const int64_t* idx, const double* scale, uint32_t n) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t < n) {
    const int64_t p0 = ptr[t], p1 = ptr[t + 1];
    double u = p0 < p1 ? scale[p0] * src[idx[p0]] : 0.0;  // one term: the plain product
    for (int64_t p = p0 + 1; p < p1; ++p) u = fma(scale[p], src[idx[p]], u);
    coeffs[t] = u;
  }
}

void launch_coeff_gather(double* coeffs, const double* src, const int64_t* ptr, const int64_t* idx,
                         const double* scale, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(coeff_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, s, coeffs, src, ptr, idx, scale, n);
}

void launch_xform_any(const XformArgs& a, uint32_t n_items, hipStream_t s) {
  if (n_items == 0) return;
  const uint32_t grid = n_items < 4096u ? n_items : 4096u;
  hipLaunchKernelGGL(xform_any_kernel, dim3(grid), dim3(256), 0, s, a, n_items);
}

void launch_xform_single(const XformArgs& a, uint32_t begin, uint32_t n, uint32_t n_flat, hipStream_t s) {
  const uint32_t total = n + n_flat;
  if (total == 0) return;
  const uint32_t grid = total < 4096u ? total : 4096u;
  hipLaunchKernelGGL(xform_single_all_kernel, dim3(grid), dim3(256), 0, s, a, begin, n, n_flat);
}

void launch_xform(const XformArgs& a, int cls, uint32_t begin, uint32_t n, hipStream_t s) {
  if (n == 0) return;
  const uint32_t grid = n < 4096u ? n : 4096u;
  switch (cls) {
    case XC_SINGLE:
      hipLaunchKernelGGL(xform_single_kernel, dim3(grid), dim3(256), 0, s, a, begin, n);
      break;
    case XC_MULTI4:
      if (multi_unroll() == 4) hipLaunchKernelGGL((xform_multi_kernel<4, 4>), dim3(grid), dim3(256), 0, s, a, begin, n);
      else hipLaunchKernelGGL((xform_multi_kernel<4, 2>), dim3(grid), dim3(256), 0, s, a, begin, n);
      break;
    case XC_MULTI16:
      hipLaunchKernelGGL((xform_multi_kernel<16>), dim3(grid), dim3(256), 0, s, a, begin, n);
      break;
    case XC_TRANSPOSE:
      hipLaunchKernelGGL(xform_transpose_kernel, dim3(grid), dim3(256), 0, s, a, begin, n);
      break;
    default:
      hipLaunchKernelGGL((xform_multi_kernel<kMaxMultiSrc>), dim3(grid), dim3(256), 0, s, a, begin, n);
      break;
  }
}

// ---------------------------------------------------------------------------
// Shadow write counts (debug): the element sets of the replay above, counted.
// A plan item covers plane elements [e_begin, e_begin + e_count) of outer
// indices [o_begin, o_begin + o_count) in every destination layout of its
// block; a transpose item is one 64 x 64 tile of one plane.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void shadow_plan_kernel(const XBlock* blocks, const XLayout* layouts,
                                                         const XItem* items, uint32_t n, uint32_t* cnt) {
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const XItem it = items[i];
    const XBlock& b = blocks[it.block];
    const int64_t n0 = b.dims[0], n1 = b.rank > 1 ? b.dims[1] : 1;
    int64_t idx[kMaxRank];
    if (b.kind == XK_TRANSPOSE) {
      const int64_t T = kTransposeTile, ntx = (n0 + T - 1) / T;
      const int64_t x0 = (int64_t)(it.e_begin % ntx) * T, y0 = (int64_t)(it.e_begin / ntx) * T;
      outer_index(b, it.o_begin, idx);
      const XLayout& L = layouts[b.lay];
      const int64_t base = base_offset(L, idx, b.rank);
      for (int e = threadIdx.x; e < T * T; e += 256) {
        const int64_t x = x0 + e % T, y = y0 + e / T;
        if (x < n0 && y < n1) atomicAdd(cnt + base + x * L.strides[0] + y * L.strides[1], 1u);
      }
      continue;
    }
    for (uint64_t o = it.o_begin; o < it.o_begin + it.o_count; ++o) {
      outer_index(b, o, idx);
      for (int j = 0; j < b.ndst; ++j) {
        const XLayout& L = layouts[b.lay + j];
        const int64_t base = base_offset(L, idx, b.rank);
        const int64_t s1 = b.rank > 1 ? L.strides[1] : 0;
        for (uint32_t e = it.e_begin + threadIdx.x; e < it.e_begin + it.e_count; e += 256) {
          const int64_t x = e % n0, y = e / n0;
          atomicAdd(cnt + base + x * L.strides[0] + y * s1, 1u);
        }
      }
    }
  }
}

// A tile: rows [tm bm, +bm) x columns [tn bn, +bn) of its job, the valid part.
// With an output map, element e of the core goes to omap[e].idx (if >= 0).
__global__ __launch_bounds__(256) void shadow_tiles_kernel(const GJob* jobs, const GTile* tiles, uint32_t n,
                                                          const TinyEntry* omap, uint32_t* cnt) {
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const GTile t = tiles[i];
    if (t.job < 0) continue;
    const GJob j = jobs[t.job];
    const int bm = tile_rows(t.cfg), bn = tile_cols(t.cfg);
    const int m0 = t.tm * bm, n0 = t.tn * bn;
    for (int e = threadIdx.x; e < bm * bn; e += 256) {
      const int row = m0 + e % bm, col = n0 + e / bm;
      if (row >= j.m || col >= j.n) continue;
      const int64_t c = j.d_off + row + (int64_t)col * j.ldd;
      const int64_t d = omap ? omap[c].idx : c;
      if (d >= 0) atomicAdd(cnt + d, 1u);
    }
  }
}

__global__ __launch_bounds__(256) void shadow_regions_kernel(const int64_t* off, const int64_t* len, uint32_t n,
                                                            uint32_t* cnt) {
  for (uint32_t r = blockIdx.x; r < n; r += gridDim.x)
    for (int64_t e = threadIdx.x; e < len[r]; e += 256) atomicAdd(cnt + off[r] + e, 1u);
}

__global__ __launch_bounds__(256) void shadow_list_kernel(const int64_t* idx, uint32_t n, uint32_t* cnt) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) atomicAdd(cnt + idx[i], 1u);
}

__global__ __launch_bounds__(256) void shadow_reduce_kernel(const uint32_t* cnt, int64_t n, unsigned long long* stats) {
  __shared__ unsigned long long part[2];
  if (threadIdx.x < 2) part[threadIdx.x] = 0ull;
  __syncthreads();
  unsigned long long sum = 0, dup = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const uint32_t c = cnt[i];
    sum += c;
    dup += c > 1u ? 1ull : 0ull;
  }
  atomicAdd(&part[0], sum);
  atomicAdd(&part[1], dup);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(stats, part[0]);
    atomicAdd(stats + 1, part[1]);
  }
}

static uint32_t shadow_grid(uint64_t n) { return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(n, 1), 4096); }

void launch_shadow_plan(const XBlock* blocks, const XLayout* layouts, const XItem* items, uint32_t n_items,
                        uint32_t* cnt, hipStream_t s) {
  if (n_items == 0) return;
  hipLaunchKernelGGL(shadow_plan_kernel, dim3(shadow_grid(n_items)), dim3(256), 0, s, blocks, layouts, items, n_items, cnt);
}

void launch_shadow_tiles(const GJob* jobs, const GTile* tiles, uint32_t n_tiles, uint32_t* cnt, hipStream_t s) {
  if (n_tiles == 0) return;
  hipLaunchKernelGGL(shadow_tiles_kernel, dim3(shadow_grid(n_tiles)), dim3(256), 0, s, jobs, tiles, n_tiles,
                     (const TinyEntry*)nullptr, cnt);
}

void launch_shadow_regions(const int64_t* off, const int64_t* len, uint32_t n, uint32_t* cnt, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(shadow_regions_kernel, dim3(shadow_grid(n)), dim3(256), 0, s, off, len, n, cnt);
}

void launch_shadow_tiny(const GJob* jobs, const GTile* tiles, uint32_t n_tiles, const TinyEntry* omap,
                        const int64_t* fix, uint32_t n_fix, uint32_t* cnt, hipStream_t s) {
  if (n_tiles) hipLaunchKernelGGL(shadow_tiles_kernel, dim3(shadow_grid(n_tiles)), dim3(256), 0, s, jobs, tiles, n_tiles, omap, cnt);
  if (n_fix) hipLaunchKernelGGL(shadow_list_kernel, dim3((n_fix + 255) / 256), dim3(256), 0, s, fix, n_fix, cnt);
}

void launch_shadow_reduce(const uint32_t* cnt, int64_t n, unsigned long long* stats, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(shadow_reduce_kernel, dim3(shadow_grid((uint64_t)(n + 255) / 256)), dim3(256), 0, s, cnt, n, stats);
}

}  // namespace tnt_dev

