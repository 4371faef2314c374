// Block GMRES of the reference (CSVector::BlockGMRES, gmres.cpp:337-1038) on
// the GPU: right preconditioning, no restart, niter block steps at most.
//
// The block Arnoldi basis is the Q factor of a Householder QR of
// [R0, A M^-1 Y_0, A M^-1 Y_1, ...] that grows by nr columns per step, as in
// the reference (UpdateQrd / MvmQBlk, globals.cpp:531-590, 778-840). Here the
// reflectors live in one column-major panel V (nl x mmax, HBM-resident, unit
// leading entries stored) together with the compact-WY factor T
// (Q = I - V T V^T, upper triangular) and its transpose, so applying Q or
// Q^T to a block costs two streaming passes over V instead of one pass per
// reflector:
//   Y_s  = Q E_s         gather rows of V, S = T (..), Y = E - V S
//   W    = A M^-1 Y_s
//   W'   = Q^T W         Z = V^T W (tall reduction), X = T^T Z, W' = W - V X
//   panel QR of W'[m:]   nr reflectors, one column at a time
//   T    extended        G = V^T V_new (tall reduction), T12 = -T11 G T22
// The block Hessenberg column (rows < m of W' plus the panel's R) is reduced
// by the previous block rotations and one new rotation (the explicit Q of a
// 2nr x 2nr Householder QR, BlockGivens / ApplyBlockGivens, gmres.cpp:30-90)
// in a one-block kernel that also tests every column's residual
// (gmres.cpp:742-750). At exit: back substitution (SolveTriangular,
// gmres.cpp:130-160), x += M^-1 Q [z; 0].
//
// All scalars stay on the device; the host polls the done flag every
// check_every steps and every kernel queued past the end returns at entry.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

constexpr int kMaxNR = 8;
enum : int { kBgDone = 0, kBgIter = 1, kBgBreak = 2, kBgConv = 3 };

__device__ __forceinline__ bool stopped(const int* done) { return done && *done; }

// Rows per lane of the tall reduction: the lane keeps RPL x NR values of W in registers.
template <int NR> constexpr int rpl() { return NR <= 2 ? 16 : NR <= 4 ? 8 : NR <= 6 ? 5 : 4; }

// out[c*ldo + i] = base(i, c) - sum_j V[j*nl + i] X[j*NR + c] for rows i < nrows,
// base(i, c) = [i == ebase + c] when base == nullptr, else base[c*ldb + i] for
// i < nbase and 0 beyond. With gram != nullptr each block also writes the
// partial sums over rows i > grow0 of out(i, 0) * out(i, l) (the first panel
// column's Gram row) to gram[block * NR + l].
template <int NR>
__global__ __launch_bounds__(256) void k_bg_apply(int64_t nrows, int64_t nl, int m, const double* __restrict__ V,
                                                  const double* __restrict__ X, const double* __restrict__ base,
                                                  int64_t ldb, int64_t nbase, int64_t ebase, double* __restrict__ out,
                                                  int64_t ldo, double* __restrict__ gram, int64_t grow0,
                                                  const int* done) {
  if (stopped(done)) return;
  constexpr int JT = 512;
  __shared__ double xs[JT * NR];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double acc[NR];
#pragma unroll
  for (int c = 0; c < NR; ++c) {
    double v = 0.0;
    if (i < nrows) {
      if (base) v = i < nbase ? base[(size_t)c * ldb + i] : 0.0;
      else v = i == ebase + c ? 1.0 : 0.0;
    }
    acc[c] = v;
  }
  for (int j0 = 0; j0 < m; j0 += JT) {
    const int jn = min(JT, m - j0);
    __syncthreads();
    for (int t = threadIdx.x; t < jn * NR; t += blockDim.x) xs[t] = X[(size_t)j0 * NR + t];
    __syncthreads();
    if (i < nrows) {
      const double* vp = V + (size_t)j0 * nl + i;
      int j = 0;
      for (; j + 8 <= jn; j += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = vp[(size_t)(j + u) * nl];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
          for (int c = 0; c < NR; ++c) acc[c] -= v[u] * xs[(j + u) * NR + c];
      }
      for (; j < jn; ++j) {
        const double v = vp[(size_t)j * nl];
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[c] -= v * xs[j * NR + c];
      }
    }
  }
  if (i < nrows) {
#pragma unroll
    for (int c = 0; c < NR; ++c) out[(size_t)c * ldo + i] = acc[c];
  }
  if (gram) {
    double g[NR];
    const bool in = i < nrows && i > grow0;
#pragma unroll
    for (int l = 0; l < NR; ++l) g[l] = in ? acc[0] * acc[l] : 0.0;
    block_partials<NR>(g, gram);
  }
}

// Tall reduction Z = V[:, 0:mcols]^T W over rows < nrows (W zero at rows >=
// nwrows): block b owns rows [b*CH, (b+1)*CH), each wave a strided subset of
// the columns; part[(j*NR + c)*nb + b] receives the block's sum.
template <int NR>
__global__ __launch_bounds__(256) void k_bg_vtw(int64_t nrows, int64_t nl, int mcols, const double* __restrict__ V,
                                                const double* __restrict__ W, int64_t ldw, int64_t nwrows,
                                                double* __restrict__ part, int nb, const int* done) {
  if (stopped(done)) return;
  constexpr int R = rpl<NR>();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int64_t r0 = (int64_t)blockIdx.x * (64 * R) + lane;
  double w[R][NR];
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int64_t r = r0 + 64 * t;
#pragma unroll
    for (int c = 0; c < NR; ++c) w[t][c] = r < nwrows ? W[(size_t)c * ldw + r] : 0.0;
  }
  for (int j = wave; j < mcols; j += nw) {
    const double* vj = V + (size_t)j * nl;
    double v[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int64_t r = r0 + 64 * t;
      v[t] = r < nrows ? vj[r] : 0.0;
    }
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) {
      double a = 0.0;
#pragma unroll
      for (int t = 0; t < R; ++t) a += v[t] * w[t][c];
      acc[c] = a;
    }
#pragma unroll
    for (int c = 0; c < NR; ++c)
      for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < NR; ++c) part[((size_t)j * NR + c) * nb + blockIdx.x] = acc[c];
    }
  }
}

// out[t] = sum_b part[t*nb + b], one wave per output.
__global__ __launch_bounds__(256) void k_bg_colsum(int M, int nb, const double* __restrict__ part,
                                                   double* __restrict__ out, const int* done) {
  if (stopped(done)) return;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= M) return;
  const double* p = part + (size_t)t * nb;
  double s = 0.0;
  for (int b = lane; b < nb; b += 64) s += p[b];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  if (lane == 0) out[t] = s;
}

// out[r*nr + c] = sum_j M(r, j) Z[j*nr + c] over rows r < m of a column-major
// triangular M (ld ldt): upper -> j in [r, m), lower -> j in [0, r].
__global__ __launch_bounds__(256) void k_bg_trmv(int m, int nr, const double* __restrict__ Mt, int64_t ldt,
                                                 int upper, const double* __restrict__ Z, double* __restrict__ out,
                                                 const int* done) {
  if (stopped(done)) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double acc[kMaxNR] = {};
  const int j0 = upper ? r : 0, j1 = upper ? m : r + 1;
  for (int j = j0; j < j1; ++j) {
    const double t = Mt[(size_t)j * ldt + r];
#pragma unroll
    for (int c = 0; c < kMaxNR; ++c)
      if (c < nr) acc[c] += t * Z[(size_t)j * nr + c];
  }
#pragma unroll
  for (int c = 0; c < kMaxNR; ++c)
    if (c < nr) out[(size_t)r * nr + c] = acc[c];
}

// out[j*nr + c] = V[j*nl + row0 + c], j < m (the rows of V^T E_s).
__global__ void k_bg_rows(int m, int nr, const double* __restrict__ V, int64_t nl, int64_t row0,
                          double* __restrict__ out, const int* done) {
  if (stopped(done)) return;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * nr) return;
  const int j = t / nr, c = t - j * nr;
  out[t] = V[(size_t)j * nl + row0 + c];
}

// Householder vector of panel column j (pivot row p = m + j): sums the Gram
// partials of the previous kernel, forms beta / tau (dlarfg convention:
// beta = -sign(alpha) sqrt(alpha^2 + |x|^2), tau = (beta - alpha) / beta,
// v = x / (alpha - beta), v(p) = 1), updates the pivot row of the later
// columns and leaves coef = {1 / (alpha - beta), f_(j+1), ..., f_(nr-1)} with
// f_l = tau v^T x_l for the row update. The last column also moves the R part
// of the block (rows < m + c of column c, and beta on its diagonal) to Hc and
// clears it in V, leaving pure reflectors.
__global__ __launch_bounds__(256) void k_bg_fin(int j, int nr, int m, int64_t nl, double* __restrict__ V,
                                                const double* __restrict__ gp, int nbp, double* __restrict__ coef,
                                                double* __restrict__ betas, double* __restrict__ tau,
                                                double* __restrict__ Hc, int64_t ldh, const int* done) {
  if (stopped(done)) return;
  __shared__ double g[kMaxNR];
  for (int l = j; l < nr; ++l) {
    const double s = sum_partials(gp, nbp, nr, l);
    if (threadIdx.x == 0) g[l] = s;
  }
  __syncthreads();
  const int64_t p = m + j;
  double* Vb = V + (size_t)m * nl;
  if (threadIdx.x == 0) {
    const double alpha = Vb[(size_t)j * nl + p];
    const double xn2 = g[j];
    double t = 0.0, beta = alpha, scale = 0.0;
    if (p < nl - 1 && xn2 != 0.0) {
      beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
      t = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    coef[0] = scale;
    for (int l = j + 1; l < nr; ++l) {
      const double f = t * (Vb[(size_t)l * nl + p] + scale * g[l]);
      Vb[(size_t)l * nl + p] -= f;
      coef[1 + l] = f;
    }
    tau[m + j] = t;
    betas[j] = beta;
    Vb[(size_t)j * nl + p] = 1.0;
  }
  if (j != nr - 1) return;
  __syncthreads();
  const int rows = m + nr;
  for (int t = threadIdx.x; t < rows * nr; t += blockDim.x) {
    const int c = t / rows, i = t - c * rows;
    double h = 0.0;
    if (i < m + c) {
      h = Vb[(size_t)c * nl + i];
      Vb[(size_t)c * nl + i] = 0.0;
    } else if (i == m + c) {
      h = betas[c];
    }
    Hc[(size_t)c * ldh + i] = h;
  }
}

// Row update of panel column j (rows > m + j): v = x_j * scale, x_l -= f_l v
// for l > j; the blocks' partial Gram row of column j + 1 (rows > m + j + 1)
// goes to gp_out.
template <int NR>
__global__ __launch_bounds__(256) void k_bg_upd(int j, int m, int64_t nl, double* __restrict__ V,
                                                const double* __restrict__ coef, double* __restrict__ gp_out,
                                                const int* done) {
  if (stopped(done)) return;
  const int64_t p = m + j;
  double* Vb = V + (size_t)m * nl;
  const double scale = coef[0];
  double f[NR];
#pragma unroll
  for (int l = 0; l < NR; ++l) f[l] = l > j ? coef[1 + l] : 0.0;
  double g[NR];
#pragma unroll
  for (int l = 0; l < NR; ++l) g[l] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = p + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += stride) {
    double x[NR];
#pragma unroll
    for (int l = 0; l < NR; ++l) x[l] = l >= j ? Vb[(size_t)l * nl + i] : 0.0;
    double v = 0.0;
#pragma unroll
    for (int l = 0; l < NR; ++l)
      if (l == j) v = x[l] * scale;
#pragma unroll
    for (int l = 0; l < NR; ++l) {
      if (l == j) Vb[(size_t)l * nl + i] = v;
      if (l > j) {
        x[l] -= f[l] * v;
        Vb[(size_t)l * nl + i] = x[l];
      }
    }
    if (i > p + 1) {
      double xn = 0.0;
#pragma unroll
      for (int l = 0; l < NR; ++l)
        if (l == j + 1) xn = x[l];
#pragma unroll
      for (int l = 0; l < NR; ++l)
        if (l > j) g[l] += xn * x[l];
    }
  }
  if (j + 1 < NR) block_partials<NR>(g, gp_out);
}

// T22 (the new nr x nr diagonal block of T) by the forward column recurrence
// of dlarft from the panel's own Gram matrix G[(m + a)*nr + c] = v_(m+a)^T v_(m+c).
__global__ void k_bg_t22(int m, int nr, double* __restrict__ T, double* __restrict__ TT, int64_t ldt,
                         const double* __restrict__ tau, const double* __restrict__ G, const int* done) {
  if (stopped(done) || threadIdx.x != 0) return;
  double t22[kMaxNR][kMaxNR] = {};
  for (int c = 0; c < nr; ++c) {
    const double tc = tau[m + c];
    double w[kMaxNR];
    for (int a = 0; a < c; ++a) w[a] = -tc * G[(size_t)(m + a) * nr + c];
    for (int a = 0; a < c; ++a) {
      double s = 0.0;
      for (int b = a; b < c; ++b) s += t22[a][b] * w[b];
      t22[a][c] = s;
    }
    t22[c][c] = tc;
  }
  for (int c = 0; c < nr; ++c)
    for (int a = 0; a < nr; ++a) {
      const double v = a <= c ? t22[a][c] : 0.0;
      T[(size_t)(m + c) * ldt + m + a] = v;
      TT[(size_t)(m + a) * ldt + m + c] = v;
    }
}

// T12 = -T11 (V_old^T V_new) T22 for rows r < m, into T and its transpose.
__global__ __launch_bounds__(256) void k_bg_t12(int m, int nr, double* __restrict__ T, double* __restrict__ TT,
                                                int64_t ldt, const double* __restrict__ G, const int* done) {
  if (stopped(done)) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  double a[kMaxNR] = {};
  for (int j = r; j < m; ++j) {
    const double t = T[(size_t)j * ldt + r];
#pragma unroll
    for (int c = 0; c < kMaxNR; ++c)
      if (c < nr) a[c] += t * G[(size_t)j * nr + c];
  }
  for (int c = 0; c < nr; ++c) {
    double s = 0.0;
    for (int q = 0; q <= c; ++q) s += a[q] * T[(size_t)(m + c) * ldt + m + q];
    T[(size_t)(m + c) * ldt + r] = -s;
    TT[(size_t)r * ldt + m + c] = -s;
  }
}

// dlarfg on a short column held in shared memory (k >= 2 entries).
__device__ void small_larfg(int k, double* alpha, double* x, double* tau) {
  double xn2 = 0.0;
  for (int i = 0; i < k - 1; ++i) xn2 += x[i] * x[i];
  *tau = 0.0;
  if (xn2 == 0.0) return;
  const double beta = -copysign(sqrt(*alpha * *alpha + xn2), *alpha);
  *tau = (beta - *alpha) / beta;
  const double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < k - 1; ++i) x[i] *= sc;
  *alpha = beta;
}

// Step s: the block column Hc ((s+2) nr rows) goes through the rotations
// 0..s-1 (a sliding 2nr-row window), the new rotation G_s is the explicit Q of
// the window's Householder QR, the finished rows go to R (column block s),
// G_s^T is applied to rows [s nr, (s+2) nr) of bloc, and each column's
// residual norm (rows [(s+1) nr, (s+2) nr) of bloc) is tested.
// sttype 0: resi < eps; 1: resi / resi0 < eps; 3: resi < eps ||b_c||.
__global__ __launch_bounds__(256) void k_bg_hess(int s, int nr, const double* __restrict__ Hc, int64_t ldh,
                                                 double* __restrict__ giv, double* __restrict__ Rm, int64_t ldr,
                                                 double* __restrict__ bloc, int64_t ldb, const double* __restrict__ r0n,
                                                 const double* __restrict__ bn, double* __restrict__ hist,
                                                 KState* ks, int niter, double eps, int sttype) {
  if (ks->st[kBgDone]) return;
  const int n2 = 2 * nr;
  __shared__ double win[2 * kMaxNR * kMaxNR], blk[2 * kMaxNR * kMaxNR], gs[4 * kMaxNR * kMaxNR], tb[kMaxNR];
  __shared__ double resi[kMaxNR];
  const int t = threadIdx.x;
  const bool own = t < n2 * nr;
  const int r = own ? t % n2 : 0, c = own ? t / n2 : 0;
  if (own) win[c * n2 + r] = Hc[(size_t)c * ldh + r];
  __syncthreads();
  for (int i = 0; i < s; ++i) {
    const double* G = giv + (size_t)i * n2 * n2;
    double v = 0.0;
    if (own)
      for (int q = 0; q < n2; ++q) v += G[(size_t)r * n2 + q] * win[c * n2 + q];
    __syncthreads();
    if (own) {
      if (r < nr) {
        Rm[(size_t)(s * nr + c) * ldr + i * nr + r] = v;
      } else {
        win[c * n2 + r - nr] = v;
        win[c * n2 + r] = Hc[(size_t)c * ldh + (i + 2) * nr + (r - nr)];
      }
    }
    __syncthreads();
  }
  // new rotation: Householder QR of the window, explicit Q
  if (own) blk[c * n2 + r] = win[c * n2 + r];
  __syncthreads();
  if (t == 0) {
    for (int i = 0; i < nr; ++i) {
      double* ci = blk + i * n2;
      for (int j = 0; j < i; ++j) {
        const double* vj = blk + j * n2;
        double sp = ci[j];
        for (int k = j + 1; k < n2; ++k) sp += ci[k] * vj[k];
        sp *= tb[j];
        ci[j] -= sp;
        for (int k = j + 1; k < n2; ++k) ci[k] -= sp * vj[k];
      }
      small_larfg(n2 - i, ci + i, ci + i + 1, tb + i);
    }
  }
  __syncthreads();
  if (t < n2) {
    double y[2 * kMaxNR];
    for (int q = 0; q < n2; ++q) y[q] = q == t ? 1.0 : 0.0;
    for (int j = nr - 1; j >= 0; --j) {
      const double* vj = blk + j * n2;
      double sp = y[j];
      for (int k = j + 1; k < n2; ++k) sp += vj[k] * y[k];
      sp = -sp * tb[j];
      y[j] += sp;
      for (int k = j + 1; k < n2; ++k) y[k] += sp * vj[k];
    }
    for (int q = 0; q < n2; ++q) {
      gs[t * n2 + q] = y[q];
      giv[(size_t)s * n2 * n2 + t * n2 + q] = y[q];
    }
  }
  __syncthreads();
  // G_s^T on the window (rows [s nr, (s+1) nr) are R's last rows) and on bloc
  double v = 0.0, bv = 0.0;
  if (own) {
    for (int q = 0; q < n2; ++q) {
      v += gs[r * n2 + q] * win[c * n2 + q];
      bv += gs[r * n2 + q] * bloc[(size_t)c * ldb + s * nr + q];
    }
  }
  __syncthreads();
  if (own) {
    if (r < nr) Rm[(size_t)(s * nr + c) * ldr + s * nr + r] = v;
    bloc[(size_t)c * ldb + s * nr + r] = bv;
  }
  __syncthreads();
  if (t < nr) {
    double ss = 0.0;
    for (int i = 0; i < nr; ++i) {
      const double b = bloc[(size_t)t * ldb + (s + 1) * nr + i];
      ss += b * b;
    }
    resi[t] = sqrt(ss);
    hist[(size_t)s * nr + t] = resi[t];
  }
  __syncthreads();
  if (t == 0) {
    int conv = 0, bad = 0;
    for (int q = 0; q < nr; ++q) {
      const double rs = resi[q];
      if (!(rs == rs)) bad = 1;
      const bool ok = sttype == 0 ? rs < eps : sttype == 1 ? rs / r0n[q] < eps : rs < eps * bn[q];
      conv += ok ? 1 : 0;
    }
    ks->st[kBgIter] = s;
    ks->st[kBgConv] = conv == nr ? 1 : 0;
    if (bad) ks->st[kBgBreak] = 1;
    if (conv == nr || bad || s + 1 >= niter) ks->st[kBgDone] = 1;
  }
}

// After the initial panel: bloc = R of the residual block (rows < nr), zero
// below; resi0 = its column norms.
__global__ void k_bg_init(int nr, const double* __restrict__ Hc, int64_t ldh, double* __restrict__ bloc,
                          int64_t ldb, double* __restrict__ r0n) {
  for (int64_t t = threadIdx.x; t < ldb * nr; t += blockDim.x) {
    const int c = (int)(t / ldb);
    const int64_t i = t - (int64_t)c * ldb;
    bloc[t] = i < nr ? Hc[(size_t)c * ldh + i] : 0.0;
  }
  __syncthreads();
  if (threadIdx.x < nr) {
    double ss = 0.0;
    for (int i = 0; i < nr; ++i) {
      const double b = Hc[(size_t)threadIdx.x * ldh + i];
      ss += b * b;
    }
    r0n[threadIdx.x] = sqrt(ss);
  }
}

__global__ void k_bg_clear(KState* ks) {
  if (threadIdx.x < 8) ks->st[threadIdx.x] = 0;
}

// Back substitution R z = bloc[0:ncolp] (column-oriented as SolveTriangular),
// z (ld ldz) zero at rows [ncolp, ldz).
__global__ __launch_bounds__(256) void k_bg_trsv(int ncolp, int nr, const double* __restrict__ Rm, int64_t ldr,
                                                 const double* __restrict__ bloc, int64_t ldb, double* __restrict__ z,
                                                 int64_t ldz) {
  __shared__ double zj[kMaxNR];
  for (int64_t t = threadIdx.x; t < ldz * nr; t += blockDim.x) {
    const int c = (int)(t / ldz);
    const int64_t i = t - (int64_t)c * ldz;
    z[t] = i < ncolp ? bloc[(size_t)c * ldb + i] : 0.0;
  }
  __syncthreads();
  for (int j = ncolp - 1; j >= 0; --j) {
    if (threadIdx.x < nr) {
      const double v = z[(size_t)threadIdx.x * ldz + j] / Rm[(size_t)j * ldr + j];
      z[(size_t)threadIdx.x * ldz + j] = v;
      zj[threadIdx.x] = v;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < j * nr; t += blockDim.x) {
      const int c = t / j, i = t - c * j;
      z[(size_t)c * ldz + i] -= Rm[(size_t)j * ldr + i] * zj[c];
    }
    __syncthreads();
  }
}

// U[j*nr + c] = sum_(i < ncolp) V[j*nl + i] z[c*ldz + i] for j < mf (V^T [z; 0]).
__global__ void k_bg_vtz(int mf, int nr, int ncolp, const double* __restrict__ V, int64_t nl,
                         const double* __restrict__ z, int64_t ldz, double* __restrict__ U) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= mf) return;
  double acc[kMaxNR] = {};
  for (int i = j; i < ncolp; ++i) {
    const double v = V[(size_t)j * nl + i];
#pragma unroll
    for (int c = 0; c < kMaxNR; ++c)
      if (c < nr) acc[c] += v * z[(size_t)c * ldz + i];
  }
  for (int c = 0; c < nr; ++c) U[(size_t)j * nr + c] = acc[c];
}

int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Launchers of the NR-templated kernels.
template <int NR>
void apply_nr(int64_t nrows, int64_t nl, int m, const double* V, const double* X, const double* base, int64_t ldb,
              int64_t nbase, int64_t ebase, double* out, int64_t ldo, double* gram, int64_t grow0, const int* done,
              hipStream_t s) {
  hipLaunchKernelGGL(k_bg_apply<NR>, dim3(std::max(1, cdiv(nrows, 256))), dim3(256), 0, s, nrows, nl, m, V, X,
                     base, ldb, nbase, ebase, out, ldo, gram, grow0, done);
}

template <int NR>
int vtw_nr(int64_t nrows, int64_t nl, int mcols, const double* V, const double* W, int64_t ldw, int64_t nwrows,
           double* part, const int* done, hipStream_t s) {
  const int nb = std::max(1, cdiv(nrows, 64 * rpl<NR>()));
  hipLaunchKernelGGL(k_bg_vtw<NR>, dim3(nb), dim3(256), 0, s, nrows, nl, mcols, V, W, ldw, nwrows, part, nb, done);
  return nb;
}

template <int NR>
void upd_nr(int j, int m, int64_t nl, double* V, const double* coef, double* gp_out, int nbp, const int* done,
            hipStream_t s) {
  hipLaunchKernelGGL(k_bg_upd<NR>, dim3(nbp), dim3(256), 0, s, j, m, nl, V, coef, gp_out, done);
}

#define BG_DISPATCH(nr, call)                       \
  switch (nr) {                                     \
    case 1: { constexpr int NRC = 1; call; } break; \
    case 2: { constexpr int NRC = 2; call; } break; \
    case 3: { constexpr int NRC = 3; call; } break; \
    case 4: { constexpr int NRC = 4; call; } break; \
    case 5: { constexpr int NRC = 5; call; } break; \
    case 6: { constexpr int NRC = 6; call; } break; \
    case 7: { constexpr int NRC = 7; call; } break; \
    default: { constexpr int NRC = 8; call; } break; \
  }

struct BgResult {
  int steps = 0, converged = 0, breakdown = 0;
};

// Block GMRES over nr (1..8) columns with nonzero initial residuals.
BgResult bgmres_core(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int nr, const double* b, double* x,
                     double* hist_host) {
  const int64_t n = A.d.n;
  const int niter = std::max(1, opt.niter);
  const int mmax = (niter + 1) * nr;
  const int64_t nl = std::max<int64_t>(n, mmax);
  const int64_t ldt = mmax, ldh = mmax + nr, ldr = (int64_t)niter * nr, ldb = (int64_t)(niter + 1) * nr;
  hipStream_t s = ctx.stream;
  double* V = ctx.vec(60, (size_t)nl * mmax * 8);
  double* T = ctx.vec(61, (size_t)ldt * mmax * 8);
  double* TT = ctx.vec(62, (size_t)ldt * mmax * 8);
  int rp = 16;
  BG_DISPATCH(nr, rp = rpl<NRC>());
  const int nbmax = std::max(1, cdiv(nl, 64 * rp));
  double* part = ctx.vec(63, (size_t)nbmax * (mmax + nr) * nr * 8);
  double* Zs = ctx.vec(64, (size_t)(mmax + nr) * nr * 8);  // reductions / gathered rows
  double* Xs = ctx.vec(65, (size_t)(mmax + nr) * nr * 8);  // T products
  double* Hc = ctx.vec(66, (size_t)ldh * nr * 8);
  double* giv = ctx.vec(67, (size_t)niter * 4 * nr * nr * 8);
  double* Rm = ctx.vec(68, (size_t)ldr * ldr * 8 + 8);
  double* bloc = ctx.vec(69, (size_t)ldb * nr * 8);
  double* tau = ctx.vec(70, (size_t)mmax * 8);
  double* small = ctx.vec(71, 64 * 8);  // coef[0..8], betas[16..24], r0n[32..40], bn[48..56]
  double* coef = small, *betas = small + 16, *r0n = small + 32, *bn = small + 48;
  double* Y = ctx.vec(72, (size_t)n * nr * 8);
  double* Z = ctx.vec(73, (size_t)n * nr * 8);
  double* W = ctx.vec(74, (size_t)n * nr * 8);
  const int gpmax = std::min(1024, std::max(1, cdiv(nl, 256)));
  double* gpa = ctx.vec(75, (size_t)std::max<int64_t>(cdiv(nl, 256), gpmax) * nr * 8);
  double* gpb = ctx.vec(76, (size_t)std::max<int64_t>(cdiv(nl, 256), gpmax) * nr * 8);
  double* hist = ctx.vec(77, (size_t)niter * nr * 8);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[kBgDone];
  const int sttype = opt.ittype == 1 ? 0 : opt.ittype == 3 ? 3 : 1;

  hipLaunchKernelGGL(k_bg_clear, dim3(1), dim3(64), 0, s, ks);
  BCM_HIP(hipGetLastError());
  if (sttype == 3) {
    const double* xs[kMaxNR];
    double* out = &ks->sc[16];
    for (int c = 0; c < nr; ++c) xs[c] = b + (size_t)c * n;
    launch_multidot((int)n, nr, xs, xs, out, ctx.rw, s);
    ctx.pull_state();
    std::vector<double> h(nr);
    for (int c = 0; c < nr; ++c) h[c] = std::sqrt(ctx.host_state->sc[16 + c]);
    BCM_HIP(hipMemcpyAsync(bn, h.data(), nr * 8, hipMemcpyHostToDevice, s));
  }

  // Panel QR of the nr columns at m (their content already in V[:, m:m+nr]),
  // the first Gram row already in gpa over nbg blocks; then T's new columns.
  auto panel = [&](int m, int nbg) {
    int nbp = nbg;
    double* gin = gpa;
    double* gout = gpb;
    for (int j = 0; j < nr; ++j) {
      hipLaunchKernelGGL(k_bg_fin, dim3(1), dim3(256), 0, s, j, nr, m, nl, V, gin, nbp, coef, betas, tau, Hc, ldh,
                         done);
      const int nbu = std::max(1, std::min(1024, cdiv(nl - (m + j + 1), 256)));
      BG_DISPATCH(nr, upd_nr<NRC>(j, m, nl, V, coef, gout, nbu, done, s));
      nbp = nbu;
      std::swap(gin, gout);
    }
    int nbv = 0;
    BG_DISPATCH(nr, nbv = vtw_nr<NRC>(nl, nl, m + nr, V, V + (size_t)m * nl, nl, nl, part, done, s));
    hipLaunchKernelGGL(k_bg_colsum, dim3(cdiv((int64_t)(m + nr) * nr, 4)), dim3(256), 0, s, (m + nr) * nr, nbv, part,
                       Zs, done);
    hipLaunchKernelGGL(k_bg_t22, dim3(1), dim3(64), 0, s, m, nr, T, TT, ldt, tau, Zs, done);
    if (m > 0) hipLaunchKernelGGL(k_bg_t12, dim3(cdiv(m, 256)), dim3(256), 0, s, m, nr, T, TT, ldt, Zs, done);
    BCM_HIP(hipGetLastError());
  };

  // R0 = b - A x into V[:, 0:nr] (zero rows beyond n), its QR, bloc and resi0
  launch_spmv(A.d, x, W, b, nr, nullptr, nullptr, ctx.rw, s);
  const int nbg0 = std::max(1, cdiv(nl, 256));
  BG_DISPATCH(nr, apply_nr<NRC>(nl, nl, 0, V, Xs, W, n, n, 0, V, nl, gpa, 0, done, s));
  panel(0, nbg0);
  hipLaunchKernelGGL(k_bg_init, dim3(1), dim3(256), 0, s, nr, Hc, ldh, bloc, ldb, r0n);
  BCM_HIP(hipGetLastError());

  const int every = std::max(1, opt.check_every);
  for (int st = 0; st < niter; ++st) {
    const int m = (st + 1) * nr;
    // Y = Q E_st
    hipLaunchKernelGGL(k_bg_rows, dim3(cdiv((int64_t)m * nr, 256)), dim3(256), 0, s, m, nr, V, nl,
                       (int64_t)st * nr, Zs, done);
    hipLaunchKernelGGL(k_bg_trmv, dim3(cdiv(m, 256)), dim3(256), 0, s, m, nr, T, ldt, 1, Zs, Xs, done);
    BG_DISPATCH(nr, apply_nr<NRC>(n, nl, m, V, Xs, nullptr, 0, 0, (int64_t)st * nr, Y, n, nullptr, 0, done, s));
    // W = A M^-1 Y
    precond_apply(F, ctx, Y, Z, opt.prec, false, nr, done);
    {
      ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, nr));
      launch_spmv(A.d, Z, W, nullptr, nr, nullptr, nullptr, ctx.rw, s, done);
    }
    // V[:, m:m+nr] = Q^T W
    int nbv = 0;
    BG_DISPATCH(nr, nbv = vtw_nr<NRC>(n, nl, m, V, W, n, n, part, done, s));
    hipLaunchKernelGGL(k_bg_colsum, dim3(cdiv((int64_t)m * nr, 4)), dim3(256), 0, s, m * nr, nbv, part, Zs, done);
    hipLaunchKernelGGL(k_bg_trmv, dim3(cdiv(m, 256)), dim3(256), 0, s, m, nr, TT, ldt, 0, Zs, Xs, done);
    BG_DISPATCH(nr, apply_nr<NRC>(nl, nl, m, V, Xs, W, n, n, 0, V + (size_t)m * nl, nl, gpa, m, done, s));
    panel(m, nbg0);
    hipLaunchKernelGGL(k_bg_hess, dim3(1), dim3(256), 0, s, st, nr, Hc, ldh, giv, Rm, ldr, bloc, ldb, r0n, bn, hist,
                       ks, niter, opt.eps, sttype);
    BCM_HIP(hipGetLastError());
    if ((st + 1) % every == 0 || st + 1 == niter) {
      ctx.pull_state();
      if (ctx.host_state->st[kBgDone]) break;
    }
  }
  ctx.pull_state();
  BgResult res;
  const int last = ctx.host_state->st[kBgIter];
  res.steps = last + 1;
  res.converged = ctx.host_state->st[kBgConv];
  res.breakdown = ctx.host_state->st[kBgBreak];
  if (hist_host) BCM_HIP(hipMemcpy(hist_host, hist, (size_t)res.steps * nr * 8, hipMemcpyDeviceToHost));

  // x += M^-1 Q [R^-1 bloc; 0]
  const int ncolp = res.steps * nr, mf = ncolp + nr;
  double* zb = ctx.vec(78, (size_t)mf * nr * 8);
  hipLaunchKernelGGL(k_bg_trsv, dim3(1), dim3(256), 0, s, ncolp, nr, Rm, ldr, bloc, ldb, zb, (int64_t)mf);
  hipLaunchKernelGGL(k_bg_vtz, dim3(cdiv(mf, 256)), dim3(256), 0, s, mf, nr, ncolp, V, nl, zb, (int64_t)mf, Zs);
  hipLaunchKernelGGL(k_bg_trmv, dim3(cdiv(mf, 256)), dim3(256), 0, s, mf, nr, T, ldt, 1, Zs, Xs, nullptr);
  BCM_HIP(hipGetLastError());
  BG_DISPATCH(nr, apply_nr<NRC>(n, nl, mf, V, Xs, zb, mf, ncolp, 0, Y, n, nullptr, 0, nullptr, s));
  precond_apply(F, ctx, Y, Z, opt.prec, false, nr);
  launch_axpby((int64_t)n * nr, 1.0, Z, 1.0, x, s);
  return res;
}

}  // namespace

void solve_block_gmres(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int nrhs, const double* b,
                       double* x, SolveStats& st, double* resi, double* hist) {
  if (nrhs < 1 || nrhs > kMaxNR) throw Error(kErrArg, "block GMRES: 1..8 right-hand sides");
  A.to_device(false);
  if (A.d.n != F.u.n()) throw Error(kErrArg, "matrix and factor sizes differ");
  const int64_t n = A.d.n;
  hipStream_t s = ctx.stream;
  KState* ks = ctx.dstate();
  // initial residual norms: columns with r0 = 0 are already solved and stay
  // out of the block (their Krylov block would be rank deficient)
  double* r = ctx.vec(79, (size_t)n * nrhs * 8);
  launch_spmv(A.d, x, r, b, nrhs, nullptr, nullptr, ctx.rw, s);
  {
    const double* xs[kMaxNR];
    for (int c = 0; c < nrhs; ++c) xs[c] = r + (size_t)c * n;
    launch_multidot((int)n, nrhs, xs, xs, &ks->sc[16], ctx.rw, s);
  }
  ctx.pull_state();
  std::vector<int> act;
  for (int c = 0; c < nrhs; ++c)
    if (ctx.host_state->sc[16 + c] > 0.0) act.push_back(c);
  const int na = (int)act.size();
  st.iterations = 0;
  st.converged = 1;
  st.breakdown = 0;
  if (na > 0) {
    const double* bb = b;
    double* xx = x;
    if (na < nrhs) {
      double* bc = ctx.vec(80, (size_t)n * na * 8);
      double* xc = ctx.vec(81, (size_t)n * na * 8);
      for (int q = 0; q < na; ++q) {
        launch_copy(n, b + (size_t)act[q] * n, bc + (size_t)q * n, s);
        launch_copy(n, x + (size_t)act[q] * n, xc + (size_t)q * n, s);
      }
      bb = bc;
      xx = xc;
    }
    std::vector<double> h;
    if (hist) h.assign((size_t)std::max(1, opt.niter) * na, 0.0);
    BgResult res = bgmres_core(A, F, ctx, opt, na, bb, xx, hist ? h.data() : nullptr);
    if (na < nrhs)
      for (int q = 0; q < na; ++q) launch_copy(n, xx + (size_t)q * n, x + (size_t)act[q] * n, s);
    st.iterations = res.steps;
    st.converged = res.converged;
    st.breakdown = res.breakdown;
    if (hist) {
      std::fill(hist, hist + (size_t)std::max(1, opt.niter) * nrhs, 0.0);
      for (int k = 0; k < res.steps; ++k)
        for (int q = 0; q < na; ++q) hist[(size_t)k * nrhs + act[q]] = h[(size_t)k * na + q];
    }
  }
  // true residuals of the result
  launch_spmv(A.d, x, r, b, nrhs, nullptr, nullptr, ctx.rw, s);
  {
    const double* xs[2 * kMaxNR];
    const double* ys[2 * kMaxNR];
    for (int c = 0; c < nrhs; ++c) {
      xs[c] = ys[c] = r + (size_t)c * n;
      xs[nrhs + c] = ys[nrhs + c] = b + (size_t)c * n;
    }
    launch_multidot((int)n, nrhs, xs, ys, &ks->sc[16], ctx.rw, s);
    launch_multidot((int)n, nrhs, xs + nrhs, ys + nrhs, &ks->sc[24], ctx.rw, s);
  }
  ctx.pull_state();
  double worst = 0.0;
  for (int c = 0; c < nrhs; ++c) {
    const double rr = std::sqrt(ctx.host_state->sc[16 + c]), bnorm = std::sqrt(ctx.host_state->sc[24 + c]);
    if (resi) resi[c] = rr;
    worst = std::max(worst, bnorm > 0 ? rr / bnorm : rr);
  }
  st.relres_true = worst;
  st.relres_rec = worst;
}

}  // namespace bcm
