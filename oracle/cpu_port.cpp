// cpu_port.cpp -- C++/OpenMP restatement of the reference's host contraction
// route, for the CPU baseline and as a second CPU checker.
//
// TEST INFRASTRUCTURE ONLY: loaded by the tests and by bench.py's CPU-baseline
// leg through oracle/cpu_port.py; the product library never links it.
//
// It replays a compiled contraction descriptor (tnt_contract_desc, the same
// structure the device backend uploads) on host cores:
//   operand / output transforms  tenet-operations/src/transform_replay.rs:3413-3606
//     Single  dst[L_d] = alpha c src[L_s] (+ beta dst)          :4428-4505
//     Multi   P_s[:, i] = src[L_s,i];  P_d = P_s U^T;  scatter    :4519-4565, :4697-4822
//     inactive destination blocks zeroed / scaled by beta        :288-330
//   core block GEMM per coupled sector (column-major jobs)       fusion_replay.rs:628-734
//   dst = alpha OUT(CORE(LHS(lhs), RHS(rhs))) + beta dst          contract/dynamic.rs:1241-1530
// Work is spread over OpenMP threads: transform blocks in element chunks,
// GEMM jobs in 192 x 192 output tiles (longest first).  The GEMM is a
// cache-blocked DGEMM of this file: packed A (192 x 256) and B (256 x 192)
// panels and an 8 x 6 register micro-kernel written with GCC vector types
// (AVX2 + FMA); no system BLAS is used.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/tenet_hip.h"

namespace {

typedef double v4d __attribute__((vector_size(32)));

constexpr int MR = 8, NR = 6;
constexpr int MC = 192, NC = 192, KC = 256;
constexpr int64_t CHUNK = 1 << 15;  // transform elements per task

// ---------------------------------------------------------------------------
// transforms
// ---------------------------------------------------------------------------
struct XTask {
  uint32_t block;  // index into desc->blocks, or n_blocks + i for inactive layout i
  int64_t begin, end;
};

inline int64_t layout_count(const tnt_layout& l) {
  int64_t n = 1;
  for (int a = 0; a < l.rank; ++a) n *= l.shape[a];
  return n;
}

// Column-major coordinates of linear index e in shape.
inline void coords(const tnt_layout& l, int64_t e, int64_t* c) {
  for (int a = 0; a < l.rank; ++a) {
    c[a] = e % l.shape[a];
    e /= l.shape[a];
  }
}

inline int64_t offset_of(const tnt_layout& l, const int64_t* c) {
  int64_t o = l.offset;
  for (int a = 0; a < l.rank; ++a) o += c[a] * l.strides[a];
  return o;
}

inline bool same_shape(const tnt_layout& a, const tnt_layout& b) {
  if (a.rank != b.rank) return false;
  for (int i = 0; i < a.rank; ++i)
    if (a.shape[i] != b.shape[i]) return false;
  return true;
}

// Visit [begin, end) of a layout's column-major element order as runs along
// axis 0: f(e0, len, c) with c the coordinates of the run's first element.
template <class F>
inline void runs(const tnt_layout& l, int64_t begin, int64_t end, F&& f) {
  int64_t c[TNT_MAX_RANK] = {0};
  if (l.rank == 0) {
    if (begin < end) f(begin, 1, c);
    return;
  }
  coords(l, begin, c);
  int64_t e = begin;
  const int64_t s0 = l.shape[0];
  while (e < end) {
    int64_t len = std::min(s0 - c[0], end - e);
    f(e, len, c);
    e += len;
    c[0] = 0;
    for (int a = 1; a < l.rank; ++a) {
      if (++c[a] < l.shape[a]) break;
      c[a] = 0;
    }
  }
}

void single_range(const tnt_transform_desc* d, const tnt_transform_block& b, const double* src,
                  double* dst, double alpha, double beta, bool overwrite, int64_t begin, int64_t end) {
  const tnt_layout& ld = d->layouts[b.dst_layout];
  const tnt_layout& ls = d->layouts[b.src_layout];
  const double k = alpha * d->coeffs[b.coeff];
  const int64_t ds0 = ld.rank ? ld.strides[0] : 0, ss0 = ls.rank ? ls.strides[0] : 0;
  runs(ld, begin, end, [&](int64_t, int64_t len, const int64_t* c) {
    double* o = dst + offset_of(ld, c);
    const double* i = src + offset_of(ls, c);
    if (overwrite) {
      for (int64_t t = 0; t < len; ++t) o[t * ds0] = k * i[t * ss0];
    } else {
      for (int64_t t = 0; t < len; ++t) o[t * ds0] = k * i[t * ss0] + beta * o[t * ds0];
    }
  });
}

void multi_range(const tnt_transform_desc* d, const tnt_transform_block& b, const double* src,
                 double* dst, double alpha, double beta, bool overwrite, int64_t begin, int64_t end) {
  const uint32_t S = b.src_count, D = b.dst_count;
  const double* U = d->coeffs + b.coeff;  // row-major U[dst][src]
  const tnt_layout* ls = d->layouts + b.src_layout;
  const tnt_layout* ld = d->layouts + b.dst_layout;
  bool uniform = true;
  for (uint32_t i = 0; i < S; ++i) uniform = uniform && same_shape(ls[i], ld[0]);
  for (uint32_t j = 0; j < D; ++j) uniform = uniform && same_shape(ld[j], ld[0]);
  std::vector<double> v(S);
  std::vector<int64_t> so(S), dof(D);
  if (uniform) {
    runs(ld[0], begin, end, [&](int64_t, int64_t len, const int64_t* c) {
      for (uint32_t i = 0; i < S; ++i) so[i] = offset_of(ls[i], c);
      for (uint32_t j = 0; j < D; ++j) dof[j] = offset_of(ld[j], c);
      for (int64_t t = 0; t < len; ++t) {
        for (uint32_t i = 0; i < S; ++i) v[i] = src[so[i] + t * (ls[i].rank ? ls[i].strides[0] : 0)];
        for (uint32_t j = 0; j < D; ++j) {
          double acc = 0.0;
          for (uint32_t i = 0; i < S; ++i) acc += U[(size_t)j * S + i] * v[i];
          double* o = dst + dof[j] + t * (ld[j].rank ? ld[j].strides[0] : 0);
          *o = overwrite ? alpha * acc : alpha * acc + beta * *o;
        }
      }
    });
    return;
  }
  // layouts of different shapes: elements correspond by linear index
  int64_t c[TNT_MAX_RANK];
  for (int64_t e = begin; e < end; ++e) {
    for (uint32_t i = 0; i < S; ++i) {
      coords(ls[i], e, c);
      v[i] = src[offset_of(ls[i], c)];
    }
    for (uint32_t j = 0; j < D; ++j) {
      double acc = 0.0;
      for (uint32_t i = 0; i < S; ++i) acc += U[(size_t)j * S + i] * v[i];
      coords(ld[j], e, c);
      double* o = dst + offset_of(ld[j], c);
      *o = overwrite ? alpha * acc : alpha * acc + beta * *o;
    }
  }
}

void inactive_range(const tnt_layout& l, double* dst, double beta, bool overwrite, int64_t begin,
                    int64_t end) {
  const int64_t s0 = l.rank ? l.strides[0] : 0;
  runs(l, begin, end, [&](int64_t, int64_t len, const int64_t* c) {
    double* o = dst + offset_of(l, c);
    for (int64_t t = 0; t < len; ++t) o[t * s0] = overwrite ? 0.0 : beta * o[t * s0];
  });
}

int transform(const tnt_transform_desc* d, const double* src, double* dst, double alpha, double beta,
              int overwrite) {
  if (!d || (!src && d->src_len) || (!dst && d->dst_len)) return TNT_E_INVALID;
  std::vector<XTask> tasks;
  auto add = [&](uint32_t id, int64_t n) {
    for (int64_t b = 0; b < n; b += CHUNK) tasks.push_back({id, b, std::min(n, b + CHUNK)});
  };
  for (size_t i = 0; i < d->n_blocks; ++i) {
    const tnt_transform_block& b = d->blocks[i];
    add((uint32_t)i, layout_count(d->layouts[b.dst_layout]));
  }
  for (size_t i = 0; i < d->n_inactive; ++i)
    add((uint32_t)(d->n_blocks + i), layout_count(d->layouts[d->inactive[i]]));
  const bool ow = overwrite != 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t t = 0; t < tasks.size(); ++t) {
    const XTask& x = tasks[t];
    if (x.block < d->n_blocks) {
      const tnt_transform_block& b = d->blocks[x.block];
      if (b.kind == TNT_BLOCK_SINGLE)
        single_range(d, b, src, dst, alpha, beta, ow, x.begin, x.end);
      else
        multi_range(d, b, src, dst, alpha, beta, ow, x.begin, x.end);
    } else {
      inactive_range(d->layouts[d->inactive[x.block - d->n_blocks]], dst, beta, ow, x.begin, x.end);
    }
  }
  return TNT_OK;
}

// ---------------------------------------------------------------------------
// GEMM
// ---------------------------------------------------------------------------
struct Operand {
  const double* p;
  int64_t ld;
  bool trans;  // op(X)(r, c) = X(c, r)
  inline double at(int64_t r, int64_t c) const { return trans ? p[c + r * ld] : p[r + c * ld]; }
};

// op(A)[i0 + 0..mc, p0 + 0..kc] -> panels of MR rows, each kc x MR (row fastest)
void pack_a(const Operand& A, int64_t i0, int64_t mc, int64_t p0, int64_t kc, double* out) {
  for (int64_t ir = 0; ir < mc; ir += MR) {
    double* o = out + ir * kc;
    const int64_t rows = std::min<int64_t>(MR, mc - ir);
    if (!A.trans && rows == MR) {
      for (int64_t p = 0; p < kc; ++p) {
        const double* col = A.p + (i0 + ir) + (p0 + p) * A.ld;
        for (int r = 0; r < MR; ++r) o[p * MR + r] = col[r];
      }
    } else {
      for (int64_t p = 0; p < kc; ++p)
        for (int r = 0; r < MR; ++r) o[p * MR + r] = r < rows ? A.at(i0 + ir + r, p0 + p) : 0.0;
    }
  }
}

// op(B)[p0 + 0..kc, j0 + 0..nc] -> panels of NR columns, each kc x NR (column fastest)
void pack_b(const Operand& B, int64_t p0, int64_t kc, int64_t j0, int64_t nc, double* out) {
  for (int64_t jr = 0; jr < nc; jr += NR) {
    double* o = out + jr * kc;
    const int64_t cols = std::min<int64_t>(NR, nc - jr);
    if (!B.trans && cols == NR) {
      const double* c0 = B.p + p0 + (j0 + jr) * B.ld;
      for (int64_t p = 0; p < kc; ++p)
        for (int c = 0; c < NR; ++c) o[p * NR + c] = c0[p + c * B.ld];
    } else {
      for (int64_t p = 0; p < kc; ++p)
        for (int c = 0; c < NR; ++c) o[p * NR + c] = c < cols ? B.at(p0 + p, j0 + jr + c) : 0.0;
    }
  }
}

// acc (column-major MR x NR) = a-panel . b-panel over kc
inline void micro_kernel(int64_t kc, const double* __restrict a, const double* __restrict b,
                         double* __restrict acc) {
  v4d c0[NR], c1[NR];
  for (int j = 0; j < NR; ++j) {
    c0[j] = v4d{0, 0, 0, 0};
    c1[j] = v4d{0, 0, 0, 0};
  }
  for (int64_t p = 0; p < kc; ++p) {
    v4d a0, a1;
    std::memcpy(&a0, a + p * MR, sizeof(a0));
    std::memcpy(&a1, a + p * MR + 4, sizeof(a1));
    const double* bp = b + p * NR;
#pragma GCC unroll 6
    for (int j = 0; j < NR; ++j) {
      c0[j] += a0 * bp[j];
      c1[j] += a1 * bp[j];
    }
  }
  for (int j = 0; j < NR; ++j) {
    std::memcpy(acc + j * MR, &c0[j], sizeof(v4d));
    std::memcpy(acc + j * MR + 4, &c1[j], sizeof(v4d));
  }
}

struct GTask {
  uint32_t job;
  int64_t i0, j0, cost;
};

struct Gemm {
  const tnt_gemm_job* jobs;
  double* dst;
  const double* lhs;
  const double* rhs;
  bool lt, rt;
  double alpha, beta;

  void tile(const GTask& t, double* ap, double* bp) const {
    const tnt_gemm_job& j = jobs[t.job];
    const int64_t m = j.rows, k = j.contracted, n = j.cols;
    Operand A{lhs + j.lhs_offset, lt ? k : m, lt};
    Operand B{rhs + j.rhs_offset, rt ? n : k, rt};
    double* Cm = dst + j.dst_offset;
    const int64_t mc = std::min<int64_t>(MC, m - t.i0), nc = std::min<int64_t>(NC, n - t.j0);
    alignas(64) double acc[MR * NR];
    for (int64_t p0 = 0; p0 < k; p0 += KC) {
      const int64_t kc = std::min<int64_t>(KC, k - p0);
      pack_a(A, t.i0, mc, p0, kc, ap);
      pack_b(B, p0, kc, t.j0, nc, bp);
      const bool first = p0 == 0;
      for (int64_t jr = 0; jr < nc; jr += NR) {
        const int64_t cols = std::min<int64_t>(NR, nc - jr);
        for (int64_t ir = 0; ir < mc; ir += MR) {
          const int64_t rows = std::min<int64_t>(MR, mc - ir);
          micro_kernel(kc, ap + ir * kc, bp + jr * kc, acc);
          for (int64_t c = 0; c < cols; ++c) {
            double* o = Cm + (t.i0 + ir) + (t.j0 + jr + c) * m;
            const double* s = acc + c * MR;
            if (first) {
              if (beta == 0.0)
                for (int64_t r = 0; r < rows; ++r) o[r] = alpha * s[r];
              else
                for (int64_t r = 0; r < rows; ++r) o[r] = alpha * s[r] + beta * o[r];
            } else {
              for (int64_t r = 0; r < rows; ++r) o[r] += alpha * s[r];
            }
          }
        }
      }
    }
  }
};

int gemm_jobs(const tnt_gemm_job* jobs, size_t n_jobs, double* dst, const double* lhs, const double* rhs,
              int lhs_op, int rhs_op, double alpha, double beta) {
  if (n_jobs && (!jobs || !dst)) return TNT_E_INVALID;
  std::vector<GTask> tasks;
  for (size_t i = 0; i < n_jobs; ++i) {
    const tnt_gemm_job& j = jobs[i];
    if (j.rows == 0 || j.cols == 0) continue;
    if (j.contracted == 0) {  // dst = beta dst: no products
      double* Cm = dst + j.dst_offset;
      const size_t len = j.rows * j.cols;
      for (size_t e = 0; e < len; ++e) Cm[e] = beta == 0.0 ? 0.0 : beta * Cm[e];
      continue;
    }
    for (int64_t i0 = 0; i0 < (int64_t)j.rows; i0 += MC)
      for (int64_t j0 = 0; j0 < (int64_t)j.cols; j0 += NC) {
        int64_t c = std::min<int64_t>(MC, j.rows - i0) * std::min<int64_t>(NC, j.cols - j0) * (int64_t)j.contracted;
        tasks.push_back({(uint32_t)i, i0, j0, c});
      }
  }
  std::stable_sort(tasks.begin(), tasks.end(), [](const GTask& a, const GTask& b) { return a.cost > b.cost; });
  Gemm g{jobs, dst, lhs, rhs, lhs_op != TNT_OP_IDENTITY, rhs_op != TNT_OP_IDENTITY, alpha, beta};
#pragma omp parallel
  {
    double* ap = static_cast<double*>(std::aligned_alloc(64, sizeof(double) * MC * KC));
    double* bp = static_cast<double*>(std::aligned_alloc(64, sizeof(double) * NC * KC));
#pragma omp for schedule(dynamic, 1)
    for (size_t t = 0; t < tasks.size(); ++t) g.tile(tasks[t], ap, bp);
    std::free(ap);
    std::free(bp);
  }
  return TNT_OK;
}

void scale_regions(const tnt_region* r, size_t n, double* dst, double beta) {
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t i = 0; i < n; ++i) {
    double* o = dst + r[i].offset;
    for (int64_t e = 0; e < r[i].len; ++e) o[e] = beta == 0.0 ? 0.0 : beta * o[e];
  }
}

}  // namespace

extern "C" {

// dst <- alpha T(src) + beta dst (overwrite != 0: dst <- alpha T(src), dst not read).
int tntc_transform(const tnt_transform_desc* d, const double* src, double* dst, double alpha, double beta,
                   int overwrite) {
  return transform(d, src, dst, alpha, beta, overwrite);
}

// Column-major jobs: dst[rows x cols] = alpha op(lhs) op(rhs) + beta dst.
int tntc_gemm_jobs(const tnt_gemm_job* jobs, size_t n_jobs, double* dst, const double* lhs,
                   const double* rhs, int lhs_op, int rhs_op, double alpha, double beta) {
  return gemm_jobs(jobs, n_jobs, dst, lhs, rhs, lhs_op, rhs_op, alpha, beta);
}

// dst = alpha OUT(CORE(LHS(lhs), RHS(rhs))) + beta dst.  Workspaces: ws_lhs
// (lhs_len, when d->lhs), ws_rhs (rhs_len, when d->rhs), ws_core (core_len,
// when d->out); unused ones may be NULL.
int tntc_contract(const tnt_contract_desc* d, double* dst, const double* lhs, const double* rhs, double alpha,
                  double beta, double* ws_lhs, double* ws_rhs, double* ws_core) {
  if (!d || !dst) return TNT_E_INVALID;
  if ((d->lhs && !ws_lhs) || (d->rhs && !ws_rhs) || (d->out && !ws_core)) return TNT_E_INVALID;
  int rc;
  const double* L = lhs;
  const double* R = rhs;
  if (d->lhs) {
    if ((rc = transform(d->lhs, lhs, ws_lhs, 1.0, 0.0, 1)) != TNT_OK) return rc;
    L = ws_lhs;
  }
  if (d->rhs) {
    if ((rc = transform(d->rhs, rhs, ws_rhs, 1.0, 0.0, 1)) != TNT_OK) return rc;
    R = ws_rhs;
  }
  if (d->out) {
    if ((rc = gemm_jobs(d->jobs, d->n_jobs, ws_core, L, R, d->lhs_op, d->rhs_op, 1.0, 0.0)) != TNT_OK) return rc;
    scale_regions(d->uncovered, d->n_uncovered, ws_core, 0.0);
    return transform(d->out, ws_core, dst, alpha, beta, beta == 0.0 ? 1 : 0);
  }
  if ((rc = gemm_jobs(d->jobs, d->n_jobs, dst, L, R, d->lhs_op, d->rhs_op, alpha, beta)) != TNT_OK) return rc;
  scale_regions(d->uncovered, d->n_uncovered, dst, beta);
  return TNT_OK;
}

int tntc_threads(void) { return omp_get_max_threads(); }

void tntc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}

}  // extern "C"
