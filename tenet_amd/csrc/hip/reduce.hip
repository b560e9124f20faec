// FP64 reductions and elementwise helpers.
//
// tnt_weighted_dot restates weighted_inner_cuda (tenet/src/typed.rs:11210-11262):
// <A,B> = sum_c dim(c) * sum_ij A_c[ij] B_c[ij].  The reference computes one
// partial per sector on the device and weights them on the host; here both
// passes run on the device with a fixed chunking and a fixed tree order, so
// the result is bitwise reproducible for a given region list (the property
// the serving loop relies on: a request's inner products do not depend on
// what else is batched with it).
#include "kernels.hpp"

namespace tnt_dev {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(256) void wdot_partials(const double* __restrict__ a,
                                                      const double* __restrict__ b,
                                                      const int64_t* __restrict__ cb,
                                                      const int64_t* __restrict__ ce,
                                                      const int32_t* __restrict__ cr,
                                                      const double* __restrict__ w,
                                                      double* __restrict__ partials) {
  __shared__ double red[4];
  const int64_t beg = cb[blockIdx.x], end = ce[blockIdx.x];
  double s = 0.0;
  for (int64_t i = beg + threadIdx.x; i < end; i += 256) s = fma(a[i], b[i], s);
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = w[cr[blockIdx.x]] * ((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ __launch_bounds__(256) void sum_partials(const double* __restrict__ p, uint32_t n,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  double s = 0.0;
  for (uint32_t i = threadIdx.x; i < n; i += 256) s += p[i];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) *out = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_weighted_dot_partials(const double* a, const double* b, const int64_t* cb,
                                  const int64_t* ce, const int32_t* cr, const double* w,
                                  double* partials, uint32_t n_chunks, hipStream_t s) {
  if (n_chunks == 0) return;
  hipLaunchKernelGGL(wdot_partials, dim3(n_chunks), dim3(256), 0, s, a, b, cb, ce, cr, w, partials);
}

void launch_sum_partials(const double* partials, uint32_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, s, partials, n, out);
}

__global__ __launch_bounds__(256) void axpby_kernel(double* __restrict__ y,
                                                     const double* __restrict__ x, int64_t n,
                                                     double alpha, double beta) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    y[i] = (beta != 0.0) ? alpha * x[i] + beta * y[i] : alpha * x[i];
}

__global__ __launch_bounds__(256) void fill_kernel(double* __restrict__ y, int64_t n, double v) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) y[i] = v;
}

__global__ __launch_bounds__(256) void zero_f32_kernel(float* __restrict__ y, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) y[i] = 0.0f;
}

__global__ __launch_bounds__(256) void scale_regions_kernel(double* __restrict__ y,
                                                             const int64_t* __restrict__ off,
                                                             const int64_t* __restrict__ len,
                                                             double beta) {
  const int64_t o = off[blockIdx.y], l = len[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < l; i += (int64_t)gridDim.x * 256)
    y[o + i] = (beta != 0.0) ? beta * y[o + i] : 0.0;
}

static uint32_t grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (uint32_t)g;
}

void launch_axpby(double* y, const double* x, int64_t n, double alpha, double beta, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, n, alpha, beta);
}

void launch_fill(double* y, int64_t n, double value, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, n, value);
}

// y = a x + b y with a = af * s[ai] (af when ai < 0) and b = bf * s[bi] (bf
// when bi < 0); bf == 0 writes a x without reading y, af == 0 writes b y
// without reading x.  Skipped entirely when gi >= 0 and s[gi] == 0 (a step a
// Krylov loop rejected on the device).  Both products are rounded before the
// sum (no fma), so an update equals the host loop's axpy (b = 1) and its
// scale-then-axpy (a = 1) bit for bit.
__global__ __launch_bounds__(256) void axpby_dev_kernel(double* y, const double* x, int64_t n,
                                                        const double* __restrict__ sc, int64_t ai, double af,
                                                        int64_t bi, double bf, int64_t gi) {
#pragma clang fp contract(off)
  if (gi >= 0 && sc[gi] == 0.0) return;
  const double a = ai >= 0 ? af * sc[ai] : af, b = bi >= 0 ? bf * sc[bi] : bf;
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    if (bf == 0.0) y[i] = a * x[i];
    else if (af == 0.0) y[i] = b * y[i];
    else y[i] = a * x[i] + b * y[i];
  }
}

void launch_axpby_dev(double* y, const double* x, int64_t n, const double* sc, int64_t ai, double af, int64_t bi,
                      double bf, int64_t gi, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(axpby_dev_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, n, sc, ai, af, bi, bf, gi);
}

// The scalar steps of a device Krylov loop, one thread, in order.
__global__ void scalars_kernel(double* __restrict__ sc, ScalarProgram prog) {
  for (int i = 0; i < prog.n; ++i) {
    const ScalarOp& o = prog.ops[i];
    const double a = sc[o.a], b = sc[o.b];
    double r = 0.0;
    switch (o.op) {
      case SCALAR_DIV: r = a / b; break;
      case SCALAR_COPY: r = a; break;
      case SCALAR_RSQRT: r = 1.0 / sqrt(a); break;
      case SCALAR_NEG: r = -a; break;
      case SCALAR_CG_GUARD: {
        // a = curvature p.Ap, b = step rho / curvature, imm != 0: check curvature
        const bool ok = isfinite(a) && (o.imm == 0.0 || a > 0.0) && isfinite(b);
        r = ok ? 1.0 : 0.0;
        break;
      }
      default: r = o.imm; break;  // SCALAR_SET
    }
    sc[o.out] = r;
  }
}

void launch_scalars(double* sc, const ScalarProgram& prog, hipStream_t s) {
  hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(1), 0, s, sc, prog);
}

void launch_zero_f32(float* y, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(zero_f32_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, n);
}

void launch_scale_regions(double* y, const int64_t* off, const int64_t* len, uint32_t n,
                          double beta, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(scale_regions_kernel, dim3(64, n), dim3(256), 0, s, y, off, len, beta);
}

// Seeded fill of the reference (tenet/src/typed.rs:2086-2092): element k of
// the flat payload receives draw k+1 of splitmix64 started at `seed`,
// mapped to ((z >> 11) / 2^52) - 1 in [-1, 1).  Every step is exact in FP64,
// so the device fill is bitwise identical to the host one.
__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void fill_unit_kernel(double* __restrict__ y, int64_t n,
                                                         uint64_t seed) {
  const int64_t stride = (int64_t)gridDim.x * 256;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    y[i] = (double)(splitmix_at(seed, (uint64_t)i) >> 11) * 0x1.0p-52 - 1.0;
}

// normal(0,1) by Box-Muller on consecutive pairs of the same stream: element
// 2p and 2p+1 use draws (2p+1, 2p+2) mapped to u in (0,1].  ln, sin and cos
// are evaluated by the fixed sequences below from +, -, *, / and sqrt only
// (every one a correctly rounded IEEE-754 double operation, contraction into
// fma disabled), so that the fill is a pure function of the seed, the same bits
// on any device and on the host: ln u = e ln2 + 2 atanh((m-1)/(m+1)) with
// u = m 2^e, m in [sqrt(1/2), sqrt(2)), series to s^23; sin / cos of
// a = (4t - rint(4t)) pi/2 in [-pi/4, pi/4] by Taylor series to a^19 / a^18,
// rotated by the quadrant rint(4t) mod 4.  Coefficients: 1/(2k+1), and
// (-1)^k / (2k+1)!, (-1)^k / (2k)!, each the double nearest the exact value.
__device__ constexpr double kAtanhC[11] = {0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
                                           0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4,
                                           0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5,
                                           0x1.8618618618618p-5, 0x1.642c8590b2164p-5};
__device__ constexpr double kSinC[9] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
                                        0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
                                        -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57};
__device__ constexpr double kCosC[9] = {-0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
                                        0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29,
                                        -0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53};
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;  // low 21 bits clear: e * kLn2Hi is exact
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;  // ln 2 - kLn2Hi
constexpr double kSqrtHalf = 0x1.6a09e667f3bcdp-1;
constexpr double kPiHalf = 0x1.921fb54442d18p+0;

// sum_k c[k] z^(k+1) as z * (c0 + z * (c1 + ...)), one rounding per operation
template <int N>
__device__ __forceinline__ double horner_z(double z, const double (&c)[N]) {
#pragma clang fp contract(off)
  double p = c[N - 1];
#pragma unroll
  for (int i = N - 2; i >= 0; --i) p = p * z + c[i];
  return p * z;
}

__device__ __forceinline__ double log_unit(double u) {
#pragma clang fp contract(off)
  int e = 0;
  double m = frexp(u, &e);
  if (m < kSqrtHalf) {
    m = m * 2.0;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f / (m + 1.0);
  const double s2 = s * 2.0;
  const double lnm = s2 + s2 * horner_z(s * s, kAtanhC);
  const double ed = (double)e;
  return ed * kLn2Hi + (lnm + ed * kLn2Lo);
}

__device__ __forceinline__ void sincos_2pi(double t, double* sn, double* cs) {
#pragma clang fp contract(off)
  const double t4 = t * 4.0;
  const double q = rint(t4);
  const double a = (t4 - q) * kPiHalf;
  const double z = a * a;
  const double sa = a + a * horner_z(z, kSinC);
  const double ca = 1.0 + horner_z(z, kCosC);
  switch ((int)q & 3) {
    case 0: *sn = sa; *cs = ca; break;
    case 1: *sn = ca; *cs = -sa; break;
    case 2: *sn = -sa; *cs = -ca; break;
    default: *sn = -ca; *cs = sa; break;
  }
}

__global__ __launch_bounds__(256) void fill_normal_kernel(double* __restrict__ y, int64_t n,
                                                           uint64_t seed) {
#pragma clang fp contract(off)
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int64_t pairs = (n + 1) / 2;
  for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < pairs; p += stride) {
    const double u1 = ((double)(splitmix_at(seed, 2 * p) >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = ((double)(splitmix_at(seed, 2 * p + 1) >> 11) + 1.0) * 0x1.0p-53;
    const double r = sqrt(log_unit(u1) * -2.0);
    double s, c;
    sincos_2pi(u2, &s, &c);
    y[2 * p] = r * c;
    if (2 * p + 1 < n) y[2 * p + 1] = r * s;
  }
}

void launch_fill_unit(double* y, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_unit_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, n, seed);
}

void launch_fill_normal(double* y, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for((n + 1) / 2)), dim3(256), 0, s, y, n, seed);
}

}  // namespace tnt_dev
