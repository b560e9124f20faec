// Complex payloads (DenseRead / DenseWrite::{C32, C64},
// tenet-dense/src/executor.rs:228-285) at the batched GEMM seam, by real
// embeddings of complex multiplication.  Each complex job becomes one real
// job of the FP64 (or FP32 / split) block GEMM -- the same 4mkn real
// multiply-adds as the four-product formula, in a single MFMA tile loop.
//
// C64 payloads with op(B) = identity and a real beta use the left-expanded
// form below (cplx_expand_lhs_kernel): only op(A) is expanded and the GEMM
// reads B and writes D in place.  Every other case uses the general form
//
//   [Cr | Ci] = [Ar | Ai] . [[Br, Bi], [-Bi, Br]]      (m x 2k) . (2k x 2n)
//
// with A = op(A_s), B = op(B_s) and (Ar, Ai), (Br, Bi), (Cr, Ci) the real and
// imaginary parts: one real m x 2k x 2n job between two HBM-bound element
// passes:
//   expand:  the stored operands, read once, written as real planes in the
//            GEMM's workspace (A: 2 doubles per element, B: 4), in the stored
//            orientation, so that a transposed or adjoint operand stays a
//            transposed operand of the real GEMM and no pass transposes;
//   combine: D = alpha (Cr + i Ci) + beta D in complex arithmetic (FP64;
//            rounded once to FP32 for C32 payloads), D not read when beta = 0.
// Plane layouts per job (column-major):
//   A identity   m x 2k, ld m :  (i, l) = Re A_s(i,l),  (i, k+l) = Im A_s(i,l)
//   A transposed 2k x m, ld 2k:  (l, i) = Re A_s(l,i),  (k+l, i) = s Im A_s(l,i)
//   B identity   2k x 2n, ld 2k: (l, j) = Re, (k+l, j) = -Im, (l, n+j) = Im, (k+l, n+j) = Re  of B_s(l,j)
//   B transposed 2n x 2k, ld 2n: (j, l) = Re, (j, k+l) = -s Im, (n+j, l) = s Im, (n+j, k+l) = Re  of B_s(j,l)
//   C            m x 2n, ld m :  (i, j) = Cr(i,j),  (i, n+j) = Ci(i,j)
// where s = -1 for the adjoint (conjugate transpose) and +1 for the transpose.
#include "kernels.hpp"

namespace tnt_dev {

template <class T>
struct Cplx {
  T re, im;
};

// Job of the first work item at or before `it` (jobs sorted by first item).
template <int Part>
__device__ __forceinline__ int cplx_job_of(const CplxJob* __restrict__ jobs, int n_jobs, uint32_t it) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const uint32_t first = Part == 0 ? jobs[mid].a_item : Part == 1 ? jobs[mid].b_item : jobs[mid].d_item;
    if (first <= it) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// One work item = kCplxItem stored elements of one operand of one job; items
// of every job's A come first, then those of every B.
template <class T>
__global__ __launch_bounds__(256) void cplx_expand_kernel(const CplxJob* __restrict__ jobs, int n_jobs,
                                                          uint32_t a_items, const Cplx<T>* __restrict__ lhs,
                                                          const Cplx<T>* __restrict__ rhs,
                                                          double* __restrict__ ws) {
  const uint32_t it = blockIdx.x;
  if (it < a_items) {
    const int j = __builtin_amdgcn_readfirstlane(cplx_job_of<0>(jobs, n_jobs, it));
    const CplxJob& J = jobs[j];
    const int64_t m = J.m, k = J.k, total = m * k;
    const int64_t e0 = (int64_t)(it - J.a_item) * kCplxItem;
    const bool tr = J.ops & CPLX_A_T;
    const double s = (J.ops & CPLX_A_CONJ) ? -1.0 : 1.0;
    double* __restrict__ P = ws + J.a_ws;
    for (int64_t e = e0 + threadIdx.x; e < e0 + kCplxItem && e < total; e += 256) {
      const Cplx<T> v = lhs[J.a_src + e];
      if (!tr) {  // stored m x k: e = i + l m
        P[e] = (double)v.re;
        P[e + total] = (double)v.im;
      } else {    // stored k x m: e = l + i k
        const int64_t l = e % k, i = e / k;
        P[l + i * 2 * k] = (double)v.re;
        P[k + l + i * 2 * k] = s * (double)v.im;
      }
    }
    return;
  }
  const uint32_t ib = it - a_items;
  const int j = __builtin_amdgcn_readfirstlane(cplx_job_of<1>(jobs, n_jobs, ib));
  const CplxJob& J = jobs[j];
  const int64_t k = J.k, n = J.n, total = k * n;
  const int64_t e0 = (int64_t)(ib - J.b_item) * kCplxItem;
  const bool tr = J.ops & CPLX_B_T;
  const double s = (J.ops & CPLX_B_CONJ) ? -1.0 : 1.0;
  double* __restrict__ P = ws + J.b_ws;
  for (int64_t e = e0 + threadIdx.x; e < e0 + kCplxItem && e < total; e += 256) {
    const Cplx<T> v = rhs[J.b_src + e];
    const double re = (double)v.re, im = (double)v.im;
    if (!tr) {  // stored k x n: e = l + j k; planes 2k x 2n, ld 2k
      const int64_t l = e % k, c = e / k;
      const int64_t ld = 2 * k;
      P[l + c * ld] = re;
      P[k + l + c * ld] = -im;
      P[l + (n + c) * ld] = im;
      P[k + l + (n + c) * ld] = re;
    } else {    // stored n x k: e = c + l n; planes 2n x 2k, ld 2n
      const int64_t c = e % n, l = e / n;
      const int64_t ld = 2 * n;
      P[c + l * ld] = re;
      P[c + (k + l) * ld] = -s * im;
      P[n + c + l * ld] = s * im;
      P[n + c + (k + l) * ld] = re;
    }
  }
}

template <class T>
__global__ __launch_bounds__(256) void cplx_combine_kernel(const CplxJob* __restrict__ jobs, int n_jobs,
                                                           const double* __restrict__ ws, Cplx<T>* __restrict__ dst,
                                                           double ar, double ai, double br, double bi) {
  const uint32_t it = blockIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(cplx_job_of<2>(jobs, n_jobs, it));
  const CplxJob& J = jobs[j];
  const int64_t total = (int64_t)J.m * J.n;
  const int64_t e0 = (int64_t)(it - J.d_item) * kCplxItem;
  const double* __restrict__ C = ws + J.c_ws;
  const bool read_d = br != 0.0 || bi != 0.0;
  for (int64_t e = e0 + threadIdx.x; e < e0 + kCplxItem && e < total; e += 256) {
    const double cr = C[e], ci = C[e + total];  // (i, j) and (i, n + j) of the m x 2n planes
    double re = ar * cr - ai * ci, im = ar * ci + ai * cr;
    Cplx<T>* p = dst + J.d_dst + e;
    if (read_d) {
      const Cplx<T> d = *p;
      const double dr = (double)d.re, di = (double)d.im;
      re += br * dr - bi * di;
      im += br * di + bi * dr;
    }
    *p = Cplx<T>{(T)re, (T)im};
  }
}

// Left-expanded form (C64 payloads, op(B) = identity, real beta): the real
// view of D (2m x n, ld 2m: rows 2i, 2i+1 = Re, Im of row i) is
//   D^ = R(alpha op(A)) . B^ + beta D^,    R(w) = [[Re w, -Im w], [Im w, Re w]]
// per element, B^ = the real view of B (2k x n, ld 2k).  Only the lhs is
// expanded (4 doubles per element), B is read and D written in place by the
// real GEMM, with no combine pass.  The expansion is stored in the stored
// orientation of A_s (R rows x C cols, planes 2R x 2C, ld 2R):
//   identity  w = alpha v,        the GEMM reads the planes as they are;
//   transpose w = conj(alpha v),  the GEMM reads them transposed (R(X)^T = R(X^H));
//   adjoint   w = conj(alpha) v,  likewise.
__global__ __launch_bounds__(256) void cplx_expand_lhs_kernel(const CplxJob* __restrict__ jobs, int n_jobs,
                                                              const Cplx<double>* __restrict__ lhs,
                                                              double* __restrict__ ws, double ar, double ai) {
  const uint32_t it = blockIdx.x;
  const int j = __builtin_amdgcn_readfirstlane(cplx_job_of<0>(jobs, n_jobs, it));
  const CplxJob& J = jobs[j];
  const int64_t total = (int64_t)J.m * J.k;
  const int64_t R = (J.ops & CPLX_A_T) ? J.k : J.m;  // stored rows
  const int64_t e0 = (int64_t)(it - J.a_item) * kCplxItem;
  const int mode = !(J.ops & CPLX_A_T) ? 0 : (J.ops & CPLX_A_CONJ) ? 2 : 1;
  double* __restrict__ P = ws + J.a_ws;
  for (int64_t e = e0 + threadIdx.x; e < e0 + kCplxItem && e < total; e += 256) {
    const Cplx<double> v = lhs[J.a_src + e];
    double wr, wi;
    if (mode == 0) {
      wr = ar * v.re - ai * v.im;
      wi = ar * v.im + ai * v.re;
    } else if (mode == 1) {
      wr = ar * v.re - ai * v.im;
      wi = -(ar * v.im + ai * v.re);
    } else {
      wr = ar * v.re + ai * v.im;
      wi = ar * v.im - ai * v.re;
    }
    // stored (p, q) = (e % R, e / R): rows 2p, 2p+1 of columns 2q and 2q+1
    const int64_t o0 = 2 * (e + (e / R) * R);
    *reinterpret_cast<double2*>(P + o0) = double2{wr, wi};
    *reinterpret_cast<double2*>(P + o0 + 2 * R) = double2{-wi, wr};
  }
}

void launch_cplx_expand_lhs(const CplxJob* jobs, int n_jobs, uint32_t a_items, const void* lhs, double* ws,
                            double ar, double ai, hipStream_t s) {
  if (!a_items) return;
  hipLaunchKernelGGL(cplx_expand_lhs_kernel, dim3(a_items), dim3(256), 0, s, jobs, n_jobs,
                     static_cast<const Cplx<double>*>(lhs), ws, ar, ai);
}

void launch_cplx_expand(const CplxJob* jobs, int n_jobs, uint32_t a_items, uint32_t b_items, const void* lhs,
                        const void* rhs, double* ws, bool c32, hipStream_t s) {
  const uint32_t items = a_items + b_items;
  if (!items) return;
  if (c32)
    hipLaunchKernelGGL(cplx_expand_kernel<float>, dim3(items), dim3(256), 0, s, jobs, n_jobs, a_items,
                       static_cast<const Cplx<float>*>(lhs), static_cast<const Cplx<float>*>(rhs), ws);
  else
    hipLaunchKernelGGL(cplx_expand_kernel<double>, dim3(items), dim3(256), 0, s, jobs, n_jobs, a_items,
                       static_cast<const Cplx<double>*>(lhs), static_cast<const Cplx<double>*>(rhs), ws);
}

void launch_cplx_combine(const CplxJob* jobs, int n_jobs, uint32_t d_items, const double* ws, void* dst,
                         double ar, double ai, double br, double bi, bool c32, hipStream_t s) {
  if (!d_items) return;
  if (c32)
    hipLaunchKernelGGL(cplx_combine_kernel<float>, dim3(d_items), dim3(256), 0, s, jobs, n_jobs, ws,
                       static_cast<Cplx<float>*>(dst), ar, ai, br, bi);
  else
    hipLaunchKernelGGL(cplx_combine_kernel<double>, dim3(d_items), dim3(256), 0, s, jobs, n_jobs, ws,
                       static_cast<Cplx<double>*>(dst), ar, ai, br, bi);
}

}  // namespace tnt_dev
