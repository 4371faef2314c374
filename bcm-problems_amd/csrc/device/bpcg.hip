// Batched preconditioned CG over many independent cells and right-hand sides.
//
// The cells' matrices and IC2 factors are uploaded as one block-diagonal
// union, so a single SpMV launch (nrhs columns) and a single pair of
// triangular sweeps serve every cell: dependency chains of different cells
// are independent and the sync-free sweep interleaves them. Each
// (cell, right-hand side) pair runs the reference's preconditioned CG
// recurrence (cg.cpp:137-289) with its own scalars; a pair that meets its
// tolerance is frozen (its x, r, p no longer change) and the launch sequence
// ends when no pair is active.
//
// Reductions are segmented by "chunks" -- runs of at most kChunk elements
// inside one cell: a chunk kernel writes one partial per (chunk, column,
// value), and one thread per pair adds its cell's chunks in a fixed order.
//
// With an FP32/BF16 factor the preconditioner is an inexact operator, so the
// recurrence uses the flexible (Polak-Ribiere) beta = (z, r - r_old) /
// (z_old, r_old), which keeps CG convergent to the FP64 tolerance.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

constexpr int kChunk = 4096;
constexpr int kPairThreads = 1024;

// Per-pair device state (struct of arrays, P = ncell * nrhs entries).
struct BPtrs {
  double *gamma, *alpha, *beta, *rr0, *bb, *pq, *relres;
  int *active, *iters, *fresh, *flag, *conv;
  int* state;  // [0] done (no active pair), [1] breakdowns
  int* cact;   // per cell: some pair of the cell is active (its rows are still computed)
};

// Refreshes the per-cell activity after a one-block kernel changed the pairs'.
__device__ __forceinline__ void cell_activity(const BPtrs& s, int npair, int nrhs) {
  __syncthreads();
  for (int c = threadIdx.x; c < npair / nrhs; c += blockDim.x) {
    int a = 0;
    for (int q = 0; q < nrhs; ++q) a |= s.active[c * nrhs + q];
    s.cact[c] = a;
  }
}

struct VecSet {
  const double* x[3];
  const double* y[3];
};

__device__ __forceinline__ double pair_thr2(const BPtrs& s, int p, int ittype, double eps2) {
  if (ittype == 0) return eps2 * s.rr0[p];
  if (ittype == 1) return eps2;
  return eps2 * s.bb[p];
}

// Sum over the chunks of cell c of value k of column q (partials laid out
// [chunk][q][k] with nrhs columns and K values).
__device__ __forceinline__ double chunk_sum(const double* partials, const int* cell_chunk0, int c, int q, int nrhs,
                                            int K, int k) {
  double s = 0.0;
  for (int ch = cell_chunk0[c]; ch < cell_chunk0[c + 1]; ++ch) s += partials[((size_t)ch * nrhs + q) * K + k];
  return s;
}

// Partial dots of K vector pairs over one chunk, for NR columns.
template <int NR, int K>
__global__ __launch_bounds__(256) void k_bp_dots(int64_t n, const int* __restrict__ lo, const int* __restrict__ hi,
                                                 VecSet v, double* partials, const int* skip) {
  if (skip && *skip) return;
  double acc[NR * K];
#pragma unroll
  for (int j = 0; j < NR * K; ++j) acc[j] = 0.0;
  const int a = lo[blockIdx.x], e = hi[blockIdx.x];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const size_t off = (size_t)q * n;
    for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
#pragma unroll
      for (int k = 0; k < K; ++k) acc[q * K + k] += v.x[k][off + i] * v.y[k][off + i];
    }
  }
  block_partials<NR * K>(acc, partials);
}

// The flexible recurrence's three dots (r.z, r.r, q.z) with z kept in FP32
// (the preconditioner's output precision; the products are formed in FP64).
// The Polak-Ribiere numerator (z, r - r_old) is -alpha (z, q), q = A p of the
// update that produced r, so r_old need not be kept.
template <int NR>
__global__ __launch_bounds__(256) void k_bp_dots_flex(int64_t n, const int* __restrict__ lo,
                                                      const int* __restrict__ hi, const double* __restrict__ r,
                                                      const double* __restrict__ qv, const float* __restrict__ z,
                                                      double* partials, const int* skip) {
  if (skip && *skip) return;
  double acc[NR * 3];
#pragma unroll
  for (int j = 0; j < NR * 3; ++j) acc[j] = 0.0;
  const int a = lo[blockIdx.x], e = hi[blockIdx.x];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const size_t off = (size_t)q * n;
    for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
      const double ri = r[off + i], zi = (double)z[off + i];
      acc[q * 3 + 0] += ri * zi;
      acc[q * 3 + 1] += ri * ri;
      acc[q * 3 + 2] += qv[off + i] * zi;
    }
  }
  block_partials<NR * 3>(acc, partials);
}

// Start of a solve: partials hold (r.r, b.b) per pair. Optional fallback to
// x0 = 0 for pairs whose initial residual is not below ||b||.
__global__ __launch_bounds__(kPairThreads) void k_bp_start(int npair, int nrhs, const int* cell_chunk0,
                                                           const double* partials, BPtrs s, double eps2, int ittype,
                                                           int fallback) {
  int any = 0;
  for (int p = threadIdx.x; p < npair; p += blockDim.x) {
    const int c = p / nrhs, q = p - c * nrhs;
    double rr = chunk_sum(partials, cell_chunk0, c, q, nrhs, 2, 0);
    const double bb = chunk_sum(partials, cell_chunk0, c, q, nrhs, 2, 1);
    s.flag[p] = 0;
    if (fallback && !(rr < bb)) {  // the guess is no better than zero
      s.flag[p] = 1;
      rr = bb;
    }
    s.bb[p] = bb;
    s.rr0[p] = rr;
    s.iters[p] = 0;
    s.fresh[p] = 1;
    s.conv[p] = 0;
    s.gamma[p] = 0.0;
    s.alpha[p] = 0.0;
    s.beta[p] = 0.0;
    const int act = rr > pair_thr2(s, p, ittype, eps2) ? 1 : 0;
    s.active[p] = act;
    if (!act) s.conv[p] = 1;
    any |= act;
  }
  cell_activity(s, npair, nrhs);
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) {
    s.state[0] = any ? 0 : 1;
    s.state[1] = 0;
  }
}

// x = 0, r = b for the pairs flagged by k_bp_start.
template <int NR>
__global__ __launch_bounds__(256) void k_bp_zero(int64_t n, const int* __restrict__ ccell, const int* __restrict__ lo,
                                                 const int* __restrict__ hi, BPtrs s, const double* __restrict__ b,
                                                 double* __restrict__ x, double* __restrict__ r) {
  const int c = ccell[blockIdx.x], a = lo[blockIdx.x], e = hi[blockIdx.x];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    if (!s.flag[c * NR + q]) continue;
    const size_t off = (size_t)q * n;
    for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
      x[off + i] = 0.0;
      r[off + i] = b[off + i];
    }
  }
}

// Block Galerkin start (no reference counterpart). Per cell, the NR guesses
// X (rows of the cell x NR) and their products A X = B - R, known from the
// initial residual R = B - A X, give G = X^T A X and F = X^T B; the guesses
// are replaced by X C with G C = F, column q of X C being the A-norm closest
// point of span(X) to the cell's solution of right-hand side q -- no farther
// from it in the A-norm than its own guess, nor than x0 = 0 -- and R by
// B - (B - R) C. This pass writes, per chunk, the partial sums of G (values
// 0 .. NR^2-1, row-major) and F (NR^2 .. 2NR^2-1).
template <int NR>
__global__ __launch_bounds__(256) void k_bp_gram(int64_t n, const int* __restrict__ lo, const int* __restrict__ hi,
                                                 const double* __restrict__ x, const double* __restrict__ b,
                                                 const double* __restrict__ r, double* partials) {
  constexpr int M = NR * NR;
  double acc[2 * M];
#pragma unroll
  for (int j = 0; j < 2 * M; ++j) acc[j] = 0.0;
  const int a = lo[blockIdx.x], e = hi[blockIdx.x];
  for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
    double xv[NR], bv[NR], av[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const size_t k = (size_t)q * n + i;
      xv[q] = x[k];
      bv[q] = b[k];
      av[q] = bv[q] - r[k];
    }
#pragma unroll
    for (int p = 0; p < NR; ++p)
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        acc[p * NR + q] += xv[p] * av[q];
        acc[M + p * NR + q] += xv[p] * bv[q];
      }
  }
  block_partials<2 * M>(acc, partials);
}

// One thread per cell: G and F from the cell's chunks, G symmetrized and
// factored by Cholesky, C = G^-1 F. A pivot not above 1e-12 of G's largest
// diagonal entry (guesses dependent to rounding, or zero) or a non-finite
// coefficient keeps the cell's guesses as they are (C = I).
template <int NR>
__global__ void k_bp_galerkin(int ncell, const int* __restrict__ cell_chunk0, const double* __restrict__ partials,
                              double* __restrict__ coef) {
  constexpr int M = NR * NR;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  double g[NR][NR], f[NR][NR];
#pragma unroll
  for (int p = 0; p < NR; ++p)
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      double sg = 0.0, sf = 0.0;
      for (int ch = cell_chunk0[c]; ch < cell_chunk0[c + 1]; ++ch) {
        sg += partials[(size_t)ch * 2 * M + p * NR + q];
        sf += partials[(size_t)ch * 2 * M + M + p * NR + q];
      }
      g[p][q] = sg;
      f[p][q] = sf;
    }
  double dmax = 0.0;
#pragma unroll
  for (int p = 0; p < NR; ++p) dmax = fmax(dmax, g[p][p]);
  bool ok = dmax > 0.0;
  double l[NR][NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    double d = g[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
    ok = ok && d > 1e-12 * dmax;
    const double ljj = ok ? sqrt(d) : 1.0;
    l[j][j] = ljj;
#pragma unroll
    for (int i = j + 1; i < NR; ++i) {
      double t = 0.5 * (g[i][j] + g[j][i]);
#pragma unroll
      for (int k = 0; k < j; ++k) t -= l[i][k] * l[j][k];
      l[i][j] = t / ljj;
    }
  }
  double cq[NR][NR];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    double y[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {  // L y = f(:, q)
      double t = f[i][q];
#pragma unroll
      for (int k = 0; k < i; ++k) t -= l[i][k] * y[k];
      y[i] = t / l[i][i];
    }
#pragma unroll
    for (int i = NR - 1; i >= 0; --i) {  // L^T c = y
      double t = y[i];
#pragma unroll
      for (int k = i + 1; k < NR; ++k) t -= l[k][i] * cq[k][q];
      cq[i][q] = t / l[i][i];
    }
  }
#pragma unroll
  for (int p = 0; p < NR; ++p)
#pragma unroll
    for (int q = 0; q < NR; ++q) ok = ok && isfinite(cq[p][q]);
#pragma unroll
  for (int p = 0; p < NR; ++p)
#pragma unroll
    for (int q = 0; q < NR; ++q) coef[(size_t)c * M + p * NR + q] = ok ? cq[p][q] : (p == q ? 1.0 : 0.0);
}

// x <- X C and r <- B - (B - R) C over the rows of one chunk.
template <int NR>
__global__ __launch_bounds__(256) void k_bp_combine(int64_t n, const int* __restrict__ ccell,
                                                    const int* __restrict__ lo, const int* __restrict__ hi,
                                                    const double* __restrict__ coef, const double* __restrict__ b,
                                                    double* __restrict__ x, double* __restrict__ r) {
  constexpr int M = NR * NR;
  __shared__ double cs[M];
  const int c = ccell[blockIdx.x], a = lo[blockIdx.x], e = hi[blockIdx.x];
  for (int k = threadIdx.x; k < M; k += blockDim.x) cs[k] = coef[(size_t)c * M + k];
  __syncthreads();
  for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
    double xv[NR], av[NR], bv[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const size_t k = (size_t)q * n + i;
      xv[q] = x[k];
      bv[q] = b[k];
      av[q] = bv[q] - r[k];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      double xn = 0.0, an = 0.0;
#pragma unroll
      for (int p = 0; p < NR; ++p) {
        xn += xv[p] * cs[p * NR + q];
        an += av[p] * cs[p * NR + q];
      }
      const size_t k = (size_t)q * n + i;
      x[k] = xn;
      r[k] = bv[q] - an;
    }
  }
}

// Convergence test and beta. Partials hold (r.z, r.r[, r_old.z]) per pair.
__global__ __launch_bounds__(kPairThreads) void k_bp_beta(int npair, int nrhs, const int* cell_chunk0,
                                                          const double* partials, int K, BPtrs s, double eps2,
                                                          int ittype, int niter) {
  if (s.state[0]) return;
  int any = 0;
  for (int p = threadIdx.x; p < npair; p += blockDim.x) {
    if (!s.active[p]) {
      s.beta[p] = 0.0;
      continue;
    }
    const int c = p / nrhs, q = p - c * nrhs;
    const double rz = chunk_sum(partials, cell_chunk0, c, q, nrhs, K, 0);
    const double rr = chunk_sum(partials, cell_chunk0, c, q, nrhs, K, 1);
    if (!(rr > pair_thr2(s, p, ittype, eps2))) {  // converged
      s.active[p] = 0;
      s.conv[p] = 1;
      continue;
    }
    if (s.iters[p] >= niter || !(rz > 0.0)) {  // budget spent, or M^-1 not positive definite
      if (!(rz > 0.0)) atomicAdd(&s.state[1], 1);
      s.active[p] = 0;
      continue;
    }
    double beta = 0.0;
    if (!s.fresh[p]) {
      // flexible: (z, r - r_old) with r - r_old = -alpha q of the last update
      const double num = K == 3 ? -s.alpha[p] * chunk_sum(partials, cell_chunk0, c, q, nrhs, K, 2) : rz;
      beta = num / s.gamma[p];
    }
    s.fresh[p] = 0;
    s.beta[p] = beta;
    s.gamma[p] = rz;
    any = 1;
  }
  cell_activity(s, npair, nrhs);
  any = __syncthreads_or(any);
  if (threadIdx.x == 0 && !any) s.state[0] = 1;
}

// p = z + beta p for active pairs.
template <int NR, class TZ>
__global__ __launch_bounds__(256) void k_bp_p(int64_t n, const int* __restrict__ ccell, const int* __restrict__ lo,
                                              const int* __restrict__ hi, BPtrs s, const TZ* __restrict__ z,
                                              double* __restrict__ pv) {
  if (s.state[0]) return;
  const int c = ccell[blockIdx.x], a = lo[blockIdx.x], e = hi[blockIdx.x];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int pr = c * NR + q;
    if (!s.active[pr]) continue;
    const double beta = s.beta[pr];
    const size_t off = (size_t)q * n;
    for (int i = a + threadIdx.x; i < e; i += blockDim.x) pv[off + i] = pv[off + i] * beta + 1.0 * (double)z[off + i];
  }
}

// alpha = gamma / (p, A p); partials hold (p.q) per pair.
__global__ __launch_bounds__(kPairThreads) void k_bp_alpha(int npair, int nrhs, const int* cell_chunk0,
                                                           const double* partials, BPtrs s) {
  if (s.state[0]) return;
  for (int p = threadIdx.x; p < npair; p += blockDim.x) {
    if (!s.active[p]) continue;
    const int c = p / nrhs, q = p - c * nrhs;
    const double pq = chunk_sum(partials, cell_chunk0, c, q, nrhs, 1, 0);
    const double alpha = s.gamma[p] / pq;
    if (!(alpha >= 0.0) || !isfinite(alpha)) {  // "Matrix is not positive definite"
      atomicAdd(&s.state[1], 1);
      s.active[p] = 0;
      s.alpha[p] = 0.0;
      continue;
    }
    s.pq[p] = pq;
    s.alpha[p] = alpha;
    s.iters[p] += 1;
  }
  cell_activity(s, npair, nrhs);
}

// x += alpha p; r -= alpha q for active pairs; with FLEX the FP32 copy of r
// that the preconditioner reads is written as well.
template <int NR, bool FLEX>
__global__ __launch_bounds__(256) void k_bp_xr(int64_t n, const int* __restrict__ ccell, const int* __restrict__ lo,
                                               const int* __restrict__ hi, BPtrs s, const double* __restrict__ pv,
                                               const double* __restrict__ qv, double* __restrict__ x,
                                               double* __restrict__ r, float* __restrict__ r32) {
  if (s.state[0]) return;
  const int c = ccell[blockIdx.x], a = lo[blockIdx.x], e = hi[blockIdx.x];
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int pr = c * NR + q;
    if (!s.active[pr]) continue;
    const double alpha = s.alpha[pr];
    const size_t off = (size_t)q * n;
    for (int i = a + threadIdx.x; i < e; i += blockDim.x) {
      const double ri = r[off + i];
      const double rn = ri + (-alpha) * qv[off + i];
      r[off + i] = rn;
      if (FLEX) r32[off + i] = (float)rn;
      x[off + i] = x[off + i] + alpha * pv[off + i];
    }
  }
}

// End of a pass: partials hold (r.r, b.b) of the true residual b - A x.
// Pairs above the tolerance that did not break down are re-armed for a
// restart from their current x (state[0] stays 0 when there is one).
__global__ __launch_bounds__(kPairThreads) void k_bp_final(int npair, int nrhs, const int* cell_chunk0,
                                                           const double* partials, BPtrs s, double eps2, int ittype,
                                                           int niter, int rearm) {
  int any = 0;
  for (int p = threadIdx.x; p < npair; p += blockDim.x) {
    const int c = p / nrhs, q = p - c * nrhs;
    const double rr = chunk_sum(partials, cell_chunk0, c, q, nrhs, 2, 0);
    const double bb = chunk_sum(partials, cell_chunk0, c, q, nrhs, 2, 1);
    s.relres[p] = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
    const bool ok = !(rr > pair_thr2(s, p, ittype, eps2));
    s.conv[p] = ok ? 1 : 0;
    if (!ok && rearm && s.iters[p] < niter && s.alpha[p] != 0.0) {
      s.active[p] = 1;
      s.fresh[p] = 1;
      any = 1;
    } else {
      s.active[p] = 0;
    }
  }
  cell_activity(s, npair, nrhs);
  any = __syncthreads_or(any);
  if (threadIdx.x == 0) s.state[0] = any ? 0 : 1;
}

__global__ void k_copy_vec(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

__global__ void k_zero_vec(int64_t n, double* __restrict__ a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a[i] = 0.0;
}

int grid_vec(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

// Work vectors of a batched solve (all n x nrhs, fp64) and the pair state.
struct BWork {
  double *x, *r, *z, *p, *q, *b, *partials;
  // profiling: algorithmic bytes per cell of the SpMV, forward and backward
  // sweep (3 x ncell, device), counted for the cells a gated launch computes
  const double* pbytes = nullptr;
  float *r32 = nullptr, *t32 = nullptr, *z32 = nullptr;  // FP32 vectors of the low-precision preconditioner
  BPtrs s;
  int* state_host;
};

BWork work(Batch& bt, Context& ctx, int nrhs, bool flex) {
  const size_t vb = (size_t)bt.n * nrhs * 8;
  const int npair = bt.ncell * nrhs;
  BWork w;
  if (flex) {
    w.r32 = reinterpret_cast<float*>(ctx.vec(82, vb / 2));
    w.t32 = reinterpret_cast<float*>(ctx.vec(83, vb / 2));
    w.z32 = reinterpret_cast<float*>(ctx.vec(84, vb / 2));
  }
  w.x = ctx.vec(40, vb);
  w.r = ctx.vec(41, vb);
  w.z = ctx.vec(42, vb);
  w.p = ctx.vec(43, vb);
  w.q = ctx.vec(44, vb);
  w.b = ctx.vec(46, vb);
  w.partials = ctx.vec(47, (size_t)bt.nchunks * nrhs * 3 * 8 + 64);
  const size_t pd = ((size_t)npair + 1) & ~size_t(1);  // keep the int arrays 8-byte aligned
  char* base = reinterpret_cast<char*>(ctx.vec(48, pd * 8 * 7 + pd * 4 * 5 + 64));
  double* d = reinterpret_cast<double*>(base);
  w.s.gamma = d;
  w.s.alpha = d + pd;
  w.s.beta = d + 2 * pd;
  w.s.rr0 = d + 3 * pd;
  w.s.bb = d + 4 * pd;
  w.s.pq = d + 5 * pd;
  w.s.relres = d + 6 * pd;
  int* iv = reinterpret_cast<int*>(d + 7 * pd);
  w.s.active = iv;
  w.s.iters = iv + pd;
  w.s.fresh = iv + 2 * pd;
  w.s.flag = iv + 3 * pd;
  w.s.conv = iv + 4 * pd;
  w.s.state = iv + 5 * pd;
  w.s.cact = reinterpret_cast<int*>(ctx.vec(85, (size_t)bt.ncell * 4 + 64));
  w.state_host = reinterpret_cast<int*>(ctx.host_state->st);
  return w;
}

template <int NR, int K>
void dots(Batch& bt, Context& ctx, const BWork& w, const double* const* xs, const double* const* ys, const int* skip) {
  VecSet v{};
  for (int k = 0; k < K; ++k) { v.x[k] = xs[k]; v.y[k] = ys[k]; }
  hipLaunchKernelGGL((k_bp_dots<NR, K>), dim3(bt.nchunks), dim3(256), 0, ctx.stream, (int64_t)bt.n,
                     bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), v, w.partials, skip);
}

// r = b - A x and the (r.r, b.b) partials.
template <int NR>
void residual_dots(Batch& bt, Context& ctx, const BWork& w) {
  launch_spmv(bt.A, w.x, w.r, w.b, NR, nullptr, nullptr, ctx.rw, ctx.stream);
  const double* xs[2] = {w.r, w.b};
  dots<NR, 2>(bt, ctx, w, xs, xs, nullptr);
}

// One PCG iteration for every active pair.
template <int NR>
void iteration(Batch& bt, Context& ctx, const BWork& w, const SolveOptions& opt, const DevTri& t1, const DevTri& t2,
               bool flex) {
  hipStream_t s = ctx.stream;
  const int* done = &w.s.state[0];
  const int npair = bt.ncell * NR;
  const double eps2 = opt.eps * opt.eps;
  // rows of cells whose pairs have all converged are left out of the sweeps and the product
  RowGate gate;
  gate.row_cell = bt.row_cell.as<int>();
  gate.active = w.s.cact;
  if (flex) {
    // FP32 sweeps straight from the FP32 copy of r (written with r) into z32
    {
      ProfScope ps(ctx, kProfSweep, w.s.cact, w.pbytes + bt.ncell, bt.ncell);
      launch_sweep(t1, w.r32, w.t32, kF32, NR, s, done, ctx.sweep_sync(), gate);
    }
    {
      ProfScope ps(ctx, kProfSweep, w.s.cact, w.pbytes + 2 * bt.ncell, bt.ncell);
      launch_sweep(t2, w.t32, w.z32, kF32, NR, s, done, ctx.sweep_sync(), gate);
    }
    hipLaunchKernelGGL(k_bp_dots_flex<NR>, dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n, bt.chunk_lo.as<int>(),
                       bt.chunk_hi.as<int>(), w.r, w.q, w.z32, w.partials, done);
  } else {
    double* tmp = ctx.vec(12, (size_t)bt.n * NR * 8);
    {
      ProfScope ps(ctx, kProfSweep, w.s.cact, w.pbytes + bt.ncell, bt.ncell);
      launch_sweep(t1, w.r, tmp, kF64, NR, s, done, ctx.sweep_sync(), gate);
    }
    {
      ProfScope ps(ctx, kProfSweep, w.s.cact, w.pbytes + 2 * bt.ncell, bt.ncell);
      launch_sweep(t2, tmp, w.z, kF64, NR, s, done, ctx.sweep_sync(), gate);
    }
    const double* xs[2] = {w.r, w.r};
    const double* ys[2] = {w.z, w.r};
    dots<NR, 2>(bt, ctx, w, xs, ys, done);
  }
  hipLaunchKernelGGL(k_bp_beta, dim3(1), dim3(kPairThreads), 0, s, npair, NR, bt.dcell_chunk0.as<int>(), w.partials,
                     flex ? 3 : 2, w.s, eps2, opt.ittype, opt.niter);
  if (flex)
    hipLaunchKernelGGL((k_bp_p<NR, float>), dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n,
                       bt.chunk_cell.as<int>(), bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), w.s,
                       (const float*)w.z32, w.p);
  else
    hipLaunchKernelGGL((k_bp_p<NR, double>), dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n,
                       bt.chunk_cell.as<int>(), bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), w.s,
                       (const double*)w.z, w.p);
  {
    ProfScope ps(ctx, kProfSpmv, w.s.cact, w.pbytes, bt.ncell);
    launch_spmv(bt.A, w.p, w.q, nullptr, NR, nullptr, nullptr, ctx.rw, s, done, gate);
  }
  {
    const double* xs[1] = {w.p};
    const double* ys[1] = {w.q};
    dots<NR, 1>(bt, ctx, w, xs, ys, done);
  }
  hipLaunchKernelGGL(k_bp_alpha, dim3(1), dim3(kPairThreads), 0, s, npair, NR, bt.dcell_chunk0.as<int>(), w.partials,
                     w.s);
  if (flex)
    hipLaunchKernelGGL((k_bp_xr<NR, true>), dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n, bt.chunk_cell.as<int>(),
                       bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), w.s, w.p, w.q, w.x, w.r, w.r32);
  else
    hipLaunchKernelGGL((k_bp_xr<NR, false>), dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n,
                       bt.chunk_cell.as<int>(), bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), w.s, w.p, w.q, w.x,
                       w.r, (float*)nullptr);
}

}  // namespace

// Captured block of check_every iterations, valid for one set of launch
// parameters and work-vector addresses.
struct BatchGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  int nrhs = 0, prec = -1, iters = 0, sweep_mode = -1, waves = 0, order = -1, flex = 0, niter = 0, ittype = -1;
  double eps = 0.0;
  std::vector<const void*> key;  // device buffers the captured launches use
  ~BatchGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
  }
};

Batch::Batch() = default;
Batch::~Batch() = default;
void Batch::drop_graph() { graph.reset(); }

void batch_build(Batch& bt, const std::vector<const Matrix*>& mats, const std::vector<const Factor*>& facts,
                 int prec_mask) {
  if (mats.empty() || mats.size() != facts.size()) throw Error(kErrArg, "batch: one factor per matrix");
  bt.ncell = (int)mats.size();
  bt.kind = facts[0]->kind;
  std::vector<BlockCsr> full(mats.size());
  std::vector<const BlockCsr*> parts(mats.size()), lower, upper;
  for (size_t c = 0; c < mats.size(); ++c) {
    if (facts[c]->kind != bt.kind) throw Error(kErrArg, "batch: mixed IC2 and ILU2 factors");
    if (facts[c]->u.n() != mats[c]->h.n()) throw Error(kErrArg, "batch: matrix and factor sizes differ");
    lower.push_back(&facts[c]->lower());
    upper.push_back(&facts[c]->u);
  }
  // upper-stored (symmetric) cells: only the upper blocks are copied and the
  // device expands them; otherwise each cell is expanded on a host thread
  bool all_symm = true;
  for (const Matrix* m : mats) all_symm = all_symm && m->symmetric;
  std::vector<int64_t> full_blocks;
  bool on_device = false;
  if (all_symm) {
    std::vector<const BlockCsr*> ups;
    for (const Matrix* m : mats) ups.push_back(&m->h);
    on_device = dev_bsr_upload_symm_parts(ups, bt.A, full_blocks);
  }
  if (!on_device) {
    parallel_for(bt.ncell, host_threads(), [&](int c) {
      if (mats[c]->symmetric) {
        full[c] = symm_mtr(mats[c]->h);
        parts[c] = &full[c];
      } else {
        parts[c] = &mats[c]->h;
      }
    });
    dev_bsr_upload_parts(parts, kF64, bt.A);
  }
  auto offdiag = [](const BlockCsr& m) {
    int64_t k = 0;
    for (int i = 0; i < m.nsup; ++i)
      for (int j = m.ia[i]; j < m.ia[i + 1]; ++j) k += m.ja[j] != i ? 1 : 0;
    return k;
  };
  bt.cell_rows.resize(mats.size());
  bt.cell_ablk.resize(mats.size());
  bt.cell_fdep.resize(mats.size());
  bt.cell_bdep.resize(mats.size());
  for (size_t c = 0; c < mats.size(); ++c) {
    bt.cell_rows[c] = mats[c]->h.nsup;
    bt.cell_ablk[c] = on_device ? full_blocks[c] : parts[c]->ia[parts[c]->nsup];
    bt.cell_fdep[c] = offdiag(*lower[c]);
    bt.cell_bdep[c] = offdiag(*upper[c]);
  }
  full = std::vector<BlockCsr>();
  bt.b = bt.A.b;
  bt.n = bt.A.n;
  bt.elem0.assign(bt.ncell + 1, 0);
  for (int c = 0; c < bt.ncell; ++c) bt.elem0[c + 1] = bt.elem0[c] + mats[c]->h.n();
  for (int p = 0; p < 3; ++p) {
    if (!(prec_mask >> p & 1)) continue;
    Factor::Image& im = bt.img[p];
    dev_tri_forward_parts(lower, p, im.fwd);
    dev_tri_backward_parts(upper, p, im.bwd);
    im.has = true;
  }
  // reduction chunks: runs of at most kChunk elements inside one cell
  std::vector<int> ccell, clo, chi;
  bt.cell_chunk0.assign(bt.ncell + 1, 0);
  for (int c = 0; c < bt.ncell; ++c) {
    for (int64_t a = bt.elem0[c]; a < bt.elem0[c + 1]; a += kChunk) {
      ccell.push_back(c);
      clo.push_back((int)a);
      chi.push_back((int)std::min<int64_t>(a + kChunk, bt.elem0[c + 1]));
    }
    bt.cell_chunk0[c + 1] = (int)ccell.size();
  }
  bt.nchunks = (int)ccell.size();
  if (bt.nchunks == 0) throw Error(kErrArg, "batch: empty cells");
  bt.chunk_cell.upload(ccell.data(), ccell.size() * 4);
  bt.chunk_lo.upload(clo.data(), clo.size() * 4);
  bt.chunk_hi.upload(chi.data(), chi.size() * 4);
  bt.dcell_chunk0.upload(bt.cell_chunk0.data(), bt.cell_chunk0.size() * 4);
  {
    std::vector<int> rc((size_t)(bt.n / bt.b));
    for (int c = 0; c < bt.ncell; ++c)
      for (int64_t r = bt.elem0[c] / bt.b; r < bt.elem0[c + 1] / bt.b; ++r) rc[(size_t)r] = c;
    bt.row_cell.upload(rc.data(), rc.size() * 4);
  }
  bt.graph.reset();
  bt.tuned = false;
}

namespace {

template <int NR>
void solve_nr(Batch& bt, Context& ctx, const SolveOptions& opt, const double* b, double* x, bool fallback,
              bool galerkin, BatchStats& st, int* iters_out, double* relres_out) {
  const Factor::Image& im = bt.img[opt.prec];
  if (!im.has) throw Error(kErrArg, "batch: factor sweeps of the requested precision were not built");
  const int npair = bt.ncell * NR;
  if (npair > 1 << 20) throw Error(kErrArg, "batch: too many (cell, right-hand side) pairs");
  hipStream_t s = ctx.stream;
  const bool flex = opt.prec != kF64;
  BWork w = work(bt, ctx, NR, flex);
  if (ctx.prof) {  // algorithmic bytes per cell of the gated launches
    const int nc = bt.ncell;
    const double B = bt.b, sv = (double)prec_bytes(opt.prec), sx = flex ? 4.0 : 8.0;
    std::vector<double> pb((size_t)3 * nc);
    for (int c = 0; c < nc; ++c) {
      const double rows = (double)bt.cell_rows[c];
      pb[c] = (double)bt.cell_ablk[c] * (B * B * 8.0 + 4.0) + rows * (4.0 + 16.0 * B * NR);
      pb[nc + c] = ((double)bt.cell_fdep[c] + rows) * (B * B * sv + 4.0) + 2.0 * sx * rows * B * NR;
      pb[2 * nc + c] = ((double)bt.cell_bdep[c] + rows) * (B * B * sv + 4.0) + 2.0 * sx * rows * B * NR;
    }
    double* d = ctx.vec(86, pb.size() * 8);
    BCM_HIP(hipMemcpyAsync(d, pb.data(), pb.size() * 8, hipMemcpyHostToDevice, s));
    BCM_HIP(hipStreamSynchronize(s));
    w.pbytes = d;
  }
  const int64_t nn = bt.n * NR;
  const double eps2 = opt.eps * opt.eps;
  const BatchTune& tu = bt.tune;
  const int saved_mode = ctx.sweep_mode;
  const int saved_waves = ctx.sync.waves_per_cu;
  const bool saved_order = ctx.sync.level_order;
  ctx.sweep_mode = tu.sweep_mode;
  ctx.sync.waves_per_cu = tu.waves_per_cu;
  ctx.sync.level_order = tu.level_order != 0;
  auto restore = [&] {
    ctx.sweep_mode = saved_mode;
    ctx.sync.waves_per_cu = saved_waves;
    ctx.sync.level_order = saved_order;
    ctx.sync.cooperative = true;
  };
  try {
    ctx.sync.ensure(im.fwd.nsup, s);
    ctx.rw.ensure();
    hipLaunchKernelGGL(k_copy_vec, dim3(grid_vec(nn)), dim3(256), 0, s, nn, b, w.b);
    hipLaunchKernelGGL(k_copy_vec, dim3(grid_vec(nn)), dim3(256), 0, s, nn, (const double*)x, w.x);
    hipLaunchKernelGGL(k_zero_vec, dim3(grid_vec(nn)), dim3(256), 0, s, nn, w.p);
    residual_dots<NR>(bt, ctx, w);
    if (galerkin) {  // each cell's guesses replaced by their A-optimal combination, then r.r and b.b again
      double* gp = ctx.vec(110, (size_t)bt.nchunks * 2 * NR * NR * 8 + 64);
      double* coef = ctx.vec(111, (size_t)bt.ncell * NR * NR * 8 + 64);
      hipLaunchKernelGGL(k_bp_gram<NR>, dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n, bt.chunk_lo.as<int>(),
                         bt.chunk_hi.as<int>(), (const double*)w.x, (const double*)w.b, (const double*)w.r, gp);
      hipLaunchKernelGGL(k_bp_galerkin<NR>, dim3((bt.ncell + 63) / 64), dim3(64), 0, s, bt.ncell,
                         bt.dcell_chunk0.as<int>(), (const double*)gp, coef);
      hipLaunchKernelGGL(k_bp_combine<NR>, dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n,
                         bt.chunk_cell.as<int>(), bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), (const double*)coef,
                         (const double*)w.b, w.x, w.r);
      const double* xs[2] = {w.r, w.b};
      dots<NR, 2>(bt, ctx, w, xs, xs, nullptr);
    }
    hipLaunchKernelGGL(k_bp_start, dim3(1), dim3(kPairThreads), 0, s, npair, NR, bt.dcell_chunk0.as<int>(),
                       w.partials, w.s, eps2, opt.ittype, fallback ? 1 : 0);
    if (fallback)
      hipLaunchKernelGGL(k_bp_zero<NR>, dim3(bt.nchunks), dim3(256), 0, s, (int64_t)bt.n, bt.chunk_cell.as<int>(),
                         bt.chunk_lo.as<int>(), bt.chunk_hi.as<int>(), w.s, w.b, w.x, w.r);
    if (flex) launch_convert(nn, w.r, kF64, w.r32, kF32, s);
    // profiling counts launches: poll after every iteration so that none is queued past the end
    const int every = ctx.prof ? 1 : std::max(1, tu.check_every);
    const bool use_graph = tu.graph && !ctx.prof;
    if (use_graph) {
      // every buffer the iteration touches exists before capture (no allocation inside it)
      std::vector<const void*> key = {w.x, w.r, w.z, w.p, w.q, w.partials, w.s.state, w.s.cact,
                                      ctx.sync.ctl.p};
      if (opt.prec == kF64) {
        key.push_back(ctx.vec(12, (size_t)nn * 8));
      } else {
        key.push_back(w.r32);
        key.push_back(w.t32);
        key.push_back(w.z32);
      }
      BatchGraph* g = bt.graph.get();
      const bool valid = g && g->nrhs == NR && g->prec == opt.prec && g->iters == every &&
                         g->sweep_mode == tu.sweep_mode && g->waves == tu.waves_per_cu &&
                         g->order == tu.level_order && g->flex == flex &&
                         g->niter == opt.niter && g->ittype == opt.ittype && g->eps == opt.eps && g->key == key;
      if (!valid) {
        bt.graph.reset(new BatchGraph());
        g = bt.graph.get();
        ctx.sync.cooperative = false;
        CaptureGuard capture;  // no set-up HIP calls of other threads until instantiation
        BCM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          for (int it = 0; it < every; ++it) iteration<NR>(bt, ctx, w, opt, im.fwd, im.bwd, flex);
        } catch (...) {
          hipGraph_t junk = nullptr;
          (void)hipStreamEndCapture(s, &junk);
          if (junk) (void)hipGraphDestroy(junk);
          throw;
        }
        BCM_HIP(hipStreamEndCapture(s, &g->g));
        BCM_HIP(hipGraphInstantiate(&g->exec, g->g, nullptr, nullptr, 0));
        ctx.sync.cooperative = true;
        g->nrhs = NR; g->prec = opt.prec; g->iters = every; g->sweep_mode = tu.sweep_mode;
        g->waves = tu.waves_per_cu; g->order = tu.level_order; g->flex = flex; g->niter = opt.niter; g->ittype = opt.ittype; g->eps = opt.eps;
        g->key = key;
      }
    }
    int passes = 0;
    while (true) {
      while (true) {
        if (use_graph) {
          BCM_HIP(hipGraphLaunch(bt.graph->exec, s));
        } else {
          for (int it = 0; it < every; ++it) iteration<NR>(bt, ctx, w, opt, im.fwd, im.bwd, flex);
        }
        BCM_HIP(hipMemcpyAsync(w.state_host, w.s.state, 8, hipMemcpyDeviceToHost, s));
        BCM_HIP(hipStreamSynchronize(s));
        if (w.state_host[0]) break;
      }
      // true residual of every pair; re-arm the ones above the tolerance
      residual_dots<NR>(bt, ctx, w);
      const int rearm = passes < opt.max_restarts ? 1 : 0;
      hipLaunchKernelGGL(k_bp_final, dim3(1), dim3(kPairThreads), 0, s, npair, NR, bt.dcell_chunk0.as<int>(),
                         w.partials, w.s, eps2, opt.ittype, opt.niter, rearm);
      BCM_HIP(hipMemcpyAsync(w.state_host, w.s.state, 8, hipMemcpyDeviceToHost, s));
      BCM_HIP(hipStreamSynchronize(s));
      if (w.state_host[0]) break;
      ++passes;
      hipLaunchKernelGGL(k_zero_vec, dim3(grid_vec(nn)), dim3(256), 0, s, nn, w.p);
      if (flex) launch_convert(nn, w.r, kF64, w.r32, kF32, s);
    }
    st.restarts = passes;
    st.breakdowns = w.state_host[1];
    hipLaunchKernelGGL(k_copy_vec, dim3(grid_vec(nn)), dim3(256), 0, s, nn, (const double*)w.x, x);
    std::vector<int> it(npair), cv(npair);
    std::vector<double> rr(npair);
    BCM_HIP(hipMemcpyAsync(it.data(), w.s.iters, npair * 4, hipMemcpyDeviceToHost, s));
    BCM_HIP(hipMemcpyAsync(cv.data(), w.s.conv, npair * 4, hipMemcpyDeviceToHost, s));
    BCM_HIP(hipMemcpyAsync(rr.data(), w.s.relres, npair * 8, hipMemcpyDeviceToHost, s));
    BCM_HIP(hipStreamSynchronize(s));
    st.pairs = npair;
    for (int p = 0; p < npair; ++p) {
      st.converged += cv[p];
      st.iters_max = std::max(st.iters_max, it[p]);
      st.iters_sum += it[p];
      st.relres_max = std::max(st.relres_max, rr[p]);
    }
    if (iters_out) std::memcpy(iters_out, it.data(), npair * 4);
    if (relres_out) std::memcpy(relres_out, rr.data(), npair * 8);
  } catch (...) {
    restore();
    throw;
  }
  restore();
}

}  // namespace

void batch_solve(Batch& bt, Context& ctx, const SolveOptions& opt, int nrhs, const double* b, double* x,
                 bool fallback, BatchStats& st, int* iters, double* relres, bool galerkin) {
  auto t0 = std::chrono::steady_clock::now();
  st = BatchStats();
  if (opt.prec < kF64 || opt.prec > kBF16) throw Error(kErrArg, "batch: unknown precision");
  if (opt.kind != kPcg && opt.kind != kBlockPcg) throw Error(kErrArg, "batch: block PCG only");
  if (bt.kind != 0) throw Error(kErrArg, "batch: block PCG needs IC2 factors");
  if (nrhs != 1 && nrhs != 3 && nrhs != 6) throw Error(kErrArg, "batch: nrhs must be 1, 3 or 6");
  auto run = [&] {
    switch (nrhs) {
      case 1: solve_nr<1>(bt, ctx, opt, b, x, fallback, galerkin, st, iters, relres); break;
      case 3: solve_nr<3>(bt, ctx, opt, b, x, fallback, galerkin, st, iters, relres); break;
      default: solve_nr<6>(bt, ctx, opt, b, x, fallback, galerkin, st, iters, relres); break;
    }
    BCM_HIP(hipStreamSynchronize(ctx.stream));
  };
  // Sync-free sweeps rely on every wave being resident; a dependency wait
  // that reaches its bound (part of the GPU held by another process) leaves
  // its sweep unfinished and raises the error flag. The solve is then
  // repeated from x0 with level-scheduled sweeps, which never wait.
  const bool syncfree = bt.tune.sweep_mode != 0;
  const int64_t nn = bt.n * nrhs;
  double* x0 = nullptr;
  if (syncfree) {
    x0 = ctx.vec(88, (size_t)nn * 8);
    launch_copy(nn, x, x0, ctx.stream);
  }
  run();
  if (syncfree && ctx.sync.error(ctx.stream)) {
    ctx.sync.clear(ctx.stream);
    launch_copy(nn, x0, x, ctx.stream);
    const BatchTune saved = bt.tune;
    bt.tune.sweep_mode = 0;
    bt.drop_graph();
    st = BatchStats();
    try {
      run();
    } catch (...) {
      bt.tune = saved;
      bt.drop_graph();
      throw;
    }
    bt.tune = saved;
    bt.drop_graph();
    st.recovered = 1;
  }
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

BatchTune batch_tune(Batch& bt, Context& ctx, const SolveOptions& opt, int nrhs, const double* b, bool allow_graph) {
  // Two stages: the sweep schedule (level-scheduled, or sync-free with a row
  // order and a launch shape) with eager launches, then graph replay and the
  // polling interval for the winner. Every candidate runs the same fixed
  // number of iterations (tolerance 0) from x0 = 0.
  const size_t vb = (size_t)bt.n * nrhs * 8;
  double* xs = ctx.vec(49, vb);
  SolveOptions o = opt;
  o.eps = 0.0;
  o.max_restarts = 0;
  o.niter = 16;  // a multiple of every polling interval tried
  const BatchTune keep = bt.tune;
  auto time_it = [&](const BatchTune& c) {
    bt.tune = c;
    bt.drop_graph();
    BatchStats st;
    BCM_HIP(hipMemsetAsync(xs, 0, vb, ctx.stream));
    batch_solve(bt, ctx, o, nrhs, b, xs, false, st, nullptr, nullptr);  // warm-up (and graph capture)
    BCM_HIP(hipMemsetAsync(xs, 0, vb, ctx.stream));
    auto t0 = std::chrono::steady_clock::now();
    batch_solve(bt, ctx, o, nrhs, b, xs, false, st, nullptr, nullptr);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return ms / std::max(1, st.iters_max);
  };
  std::vector<BatchTune> stage1;
  {
    BatchTune t;
    t.graph = 0;
    t.check_every = 16;
    t.sweep_mode = 0;
    stage1.push_back(t);
    t.sweep_mode = 1;
    for (int order = 0; order <= 1; ++order)
      for (int waves : {4, 8, 16}) {
        t.level_order = order;
        t.waves_per_cu = waves;
        stage1.push_back(t);
      }
  }
  BatchTune best = keep;
  best.ms_per_iter = 1e300;
  for (const BatchTune& c : stage1) {
    const double per = time_it(c);
    if (per < best.ms_per_iter) { best = c; best.ms_per_iter = per; }
  }
  const BatchTune sched = best;
  for (int graph = 0; graph <= (allow_graph ? 1 : 0); ++graph)
    for (int every : {4, 16}) {
      if (graph == 0 && every == 16) continue;  // measured in stage one
      BatchTune c = sched;
      c.graph = graph;
      c.check_every = every;
      const double per = time_it(c);
      if (per < best.ms_per_iter) { best = c; best.ms_per_iter = per; }
    }
  bt.tune = best;
  bt.tuned = true;
  bt.drop_graph();
  return best;
}

}  // namespace bcm
