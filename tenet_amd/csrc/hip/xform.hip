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
#include "kernels.hpp"

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

__device__ void xform_item(const XformArgs& a, const XItem it);

// Persistent grid-stride loop over work items: concurrently resident
// workgroups process consecutive items, so the waves of the whole GPU sweep
// one contiguous window of the payloads at a time (DRAM locality), like a
// plain grid-stride copy.
__global__ __launch_bounds__(256) void xform_kernel(XformArgs a, uint32_t n_items) {
  for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) xform_item(a, a.items[i]);
}

__device__ void xform_item(const XformArgs& a, const XItem it) {
  const XBlock& b = a.blocks[it.block];
  const int rank = b.rank;
  const int kind = b.kind;
  const int64_t n0 = b.dims[0];
  const int64_t n1 = rank > 1 ? b.dims[1] : 1;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;

  if (it.o_count == 1 && kind == XK_SINGLE) {
    // one outer index, up to kSingleItem plane elements: element
    // e = e_begin + 64 * (4u + wave) + lane, u < kSingleUnroll -- the whole
    // workgroup sweeps the item front to back in 256-element steps, with all
    // kSingleUnroll loads of a lane in flight before its stores
    int64_t idx[kMaxRank];
    {
      uint64_t rr = it.o_begin;
#pragma unroll
      for (int d = 2; d < kMaxRank; ++d) {
        if (d < rank) {
          idx[d] = (int64_t)(rr % (uint64_t)b.dims[d]);
          rr /= (uint64_t)b.dims[d];
        } else {
          idx[d] = 0;
        }
      }
    }
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
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u) {
      const uint32_t f = f0 + 256u * u;
      if (f < it.e_count) {
        v[u] = a.src[os + x * s0 + y * s1];
        dp[u] = od + x * d0 + y * d1;
      }
      x += sx;
      y += sy;
      if (x >= n0u) {
        x -= n0u;
        y += 1;
      }
    }
#pragma unroll
    for (int u = 0; u < kSingleUnroll; ++u) {
      const uint32_t f = f0 + 256u * u;
      if (f < it.e_count) {
        double* p = a.dst + dp[u];
        *p = readd ? c * v[u] + a.beta * *p : c * v[u];
      }
    }
    return;
  }

  // Waves split outer indices when there are enough of them, else the plane.
  uint64_t o_first, o_step;
  uint32_t e_lo, e_hi;
  if (it.o_count >= 4) {
    o_first = it.o_begin + wave;
    o_step = 4;
    e_lo = it.e_begin;
    e_hi = it.e_begin + it.e_count;
  } else {
    o_first = it.o_begin;
    o_step = 1;
    uint32_t q = (it.e_count + 3) / 4;
    e_lo = it.e_begin + min(it.e_count, q * wave);
    e_hi = it.e_begin + min(it.e_count, q * (wave + 1));
  }
  const uint64_t o_end = it.o_begin + it.o_count;
  // incremental plane walk: e -> (x, y) with e = x + n0 * y; step 64
  const int64_t dx = 64 % n0, dy = 64 / n0;
  const double alpha = a.alpha, beta = a.beta;
  const bool ow = a.overwrite != 0;
  const bool readd = !ow && beta != 0.0;

  for (uint64_t o = o_first; o < o_end; o += o_step) {
    int64_t idx[kMaxRank];
    {
      uint64_t rr = o;
#pragma unroll
      for (int d = 2; d < kMaxRank; ++d) {
        if (d < rank) {
          idx[d] = (int64_t)(rr % (uint64_t)b.dims[d]);
          rr /= (uint64_t)b.dims[d];
        } else {
          idx[d] = 0;
        }
      }
    }
    uint32_t e = e_lo + lane;
    int64_t x = e % n0, y = e / n0;
    if (kind == XK_SINGLE) {
      const XLayout& Ld = a.layouts[b.lay];
      const XLayout& Ls = a.layouts[b.lay + 1];
      const int64_t od = base_offset(Ld, idx, rank), os = base_offset(Ls, idx, rank);
      const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
      const int64_t s0 = Ls.strides[0], s1 = rank > 1 ? Ls.strides[1] : 0;
      const double c = alpha * a.coeffs[b.coeff];
      const double* __restrict__ src = a.src;
      double* __restrict__ dst = a.dst;
      // four independent loads in flight per lane before their stores
      while (e < e_hi) {
        double v[kUnroll];
        int64_t dp[kUnroll];
        int cnt = 0;
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (e < e_hi) {
            v[u] = src[os + x * s0 + y * s1];
            dp[u] = od + x * d0 + y * d1;
            cnt = u + 1;
            e += 64;
            x += dx;
            y += dy;
            if (x >= n0) { x -= n0; y += 1; }
          }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (u < cnt) {
            double* p = dst + dp[u];
            *p = readd ? c * v[u] + beta * *p : c * v[u];
          }
        }
      }
    } else if (kind == XK_MULTI) {
      const int nd = b.ndst, ns = b.nsrc;
      const XLayout* L = a.layouts + b.lay;
      const double* U = a.coeffs + b.coeff;
      for (; e < e_hi; e += 64) {
        double v[kMaxMultiSrc];
#pragma unroll
        for (int i = 0; i < kMaxMultiSrc; ++i) {
          if (i < ns) {
            const XLayout& Ls = L[nd + i];
            const int64_t off = base_offset(Ls, idx, rank) + x * Ls.strides[0] +
                                (rank > 1 ? y * Ls.strides[1] : 0);
            v[i] = a.src[off];
          } else {
            v[i] = 0.0;
          }
        }
        for (int j = 0; j < nd; ++j) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < kMaxMultiSrc; ++i)
            if (i < ns) acc = fma(U[j * ns + i], v[i], acc);
          const XLayout& Ld = L[j];
          double* p = a.dst + base_offset(Ld, idx, rank) + x * Ld.strides[0] +
                      (rank > 1 ? y * Ld.strides[1] : 0);
          *p = readd ? alpha * acc + beta * *p : alpha * acc;
        }
        x += dx;
        y += dy;
        if (x >= n0) { x -= n0; y += 1; }
      }
    } else {  // XK_SCALE: inactive destination block
      const XLayout& Ld = a.layouts[b.lay];
      const int64_t od = base_offset(Ld, idx, rank);
      const int64_t d0 = Ld.strides[0], d1 = rank > 1 ? Ld.strides[1] : 0;
      for (; e < e_hi; e += 64) {
        double* p = a.dst + od + x * d0 + y * d1;
        *p = readd ? beta * *p : 0.0;
        x += dx;
        y += dy;
        if (x >= n0) { x -= n0; y += 1; }
      }
    }
  }
  (void)n1;
}

void launch_xform(const XformArgs& a, uint32_t n_items, hipStream_t s) {
  if (n_items == 0) return;
  const uint32_t grid = n_items < 4096u ? n_items : 4096u;
  hipLaunchKernelGGL(xform_kernel, dim3(grid), dim3(256), 0, s, a, n_items);
}

}  // namespace tnt_dev
