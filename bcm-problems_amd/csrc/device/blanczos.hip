// Block Lanczos of the reference (CSVector::BlockLanczosRight / Left /
// Central, lanczos.cpp:932-3225) on the GPU: the block Golub-Kahan
// bidiagonalization of a preconditioned operator with block Givens rotations
// (an LSQR over nb right-hand sides at once), FP64 throughout.
//
// Operator of each variant (L by columns, U by rows, diagonal blocks
// pre-inverted; the transposed sweeps run on the factor images stored for
// M^-T):
//   Right    Ahat = A U^-1 L^-1, x = x0 + U^-1 L^-1 s   (lanczos.cpp:932-1718)
//   Left     Ahat = U^-1 L^-1 A, x = x0 + s             (lanczos.cpp:1719-2466)
//   Central  Ahat = L^-1 A U^-1, x = x0 + U^-1 s        (lanczos.cpp:2467-3225)
// Per iteration, as the reference:
//   Q = Ahat V - U alpha,   U beta  = QR(Q)
//   P = Ahat^T U - V beta^T, V alpha^T = QR(P)
//   block rotation [C S; -Sbar Cbar] of [rhobar; beta] (explicit Q of its
//   2nb x 2nb QR), theta = S alpha, rhobar = Cbar alpha, phi = C phibar,
//   phibar = -Sbar phibar, s += W rho^-1 phi, W = V - W rho^-1 theta.
// Every ichk iterations the true residual b - A x_k of every column is
// tested against its initial norm (Right: also ||Ahat^T r||), and the
// solve stops when every column passes (lanczos.cpp:1536-1585).
//
// The tall blocks are factored by CholeskyQR2 (two passes of Gram matrix +
// Cholesky + triangular solve, each pass a reduction and a streaming
// multiply over the n x nb block) instead of the reference's column-by-column
// Householder QR (QrdBlk): Q and R agree up to the signs of R's rows, which
// leave the iterates unchanged (the block bidiagonalization is unique up to
// an orthogonal change of basis inside each block), and the two passes keep
// Q orthonormal to rounding. All nb x nb and 2nb x 2nb work (Cholesky,
// rotation, the inverse of rho) runs in one-block kernels on the device; the
// host reads the state once per residual test.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

constexpr int kMaxNB = 8;
constexpr int kGramMax = kMaxNB * (kMaxNB + 1) / 2;
enum : int { kLzDone = 0, kLzIter = 1, kLzBreak = 2, kLzConv = 3 };

// Slots of the small state (doubles): nb x nb matrices column-major.
enum : int {
  kSAlpha = 0, kSBeta = 64, kSBetaT = 128, kSRhobar = 192, kSPhibar = 256, kSCoef1 = 320, kSCoef2 = 384,
  kSR1 = 448, kSRinv = 512, kSResi0 = 576, kSResort0 = 584, kSResi = 592, kSResort = 600, kSTotal = 608,
};

__device__ __forceinline__ bool stopped(const int* done) { return done && *done; }

int red_blocks(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)); }

// Block partial sums of the Gram matrix W^T W (upper triangle, column by
// column: index c(c+1)/2 + a for a <= c).
template <int NB>
__global__ __launch_bounds__(256) void k_lz_gram(int64_t n, const double* __restrict__ W, double* part,
                                                 const int* done) {
  if (stopped(done)) return;
  constexpr int K = NB * (NB + 1) / 2;
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double w[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) w[c] = W[(size_t)c * n + i];
#pragma unroll
    for (int c = 0; c < NB; ++c)
#pragma unroll
      for (int a = 0; a <= c; ++a) acc[c * (c + 1) / 2 + a] += w[a] * w[c];
  }
  block_partials<K>(acc, part);
}

// out = base + s X C row by row (each thread reads its row of X before
// writing, so out may alias X or base); C is nb x nb column-major in the
// small state; base may be null (zero).
template <int NB>
__global__ __launch_bounds__(256) void k_lz_mul(int64_t n, const double* X, const double* __restrict__ C,
                                                const double* base, double s, double* out, const int* done) {
  if (stopped(done)) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double x[NB], y[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) x[j] = X[(size_t)j * n + i];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      double a = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) a += x[j] * C[c * NB + j];
      y[c] = (base ? base[(size_t)c * n + i] : 0.0) + s * a;
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) out[(size_t)c * n + i] = y[c];
  }
}

// s += W coef1; W = V - W coef2 (lanczos.cpp:1486-1510), row by row.
template <int NB>
__global__ __launch_bounds__(256) void k_lz_wupd(int64_t n, double* __restrict__ W, const double* __restrict__ V,
                                                 double* __restrict__ S, const double* __restrict__ sm,
                                                 const int* done) {
  if (stopped(done)) return;
  const double* c1 = sm + kSCoef1;
  const double* c2 = sm + kSCoef2;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double w[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) w[j] = W[(size_t)j * n + i];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        a += w[j] * c1[c * NB + j];
        b += w[j] * c2[c * NB + j];
      }
      S[(size_t)c * n + i] += a;
      W[(size_t)c * n + i] = V[(size_t)c * n + i] - b;
    }
  }
}

// Upper Cholesky factor R (G = R^T R) of the nb x nb Gram matrix g and its
// inverse; false when a pivot is not positive beyond rounding (rank-deficient block).
__device__ bool chol_inv(int nb, const double* g, double* R, double* Ri) {
  for (int i = 0; i < nb * nb; ++i) R[i] = Ri[i] = 0.0;
  for (int j = 0; j < nb; ++j) {
    double d = g[j * nb + j];
    for (int k = 0; k < j; ++k) d -= R[j * nb + k] * R[j * nb + k];
    if (!(d > 1e-28 * g[j * nb + j]) || !(d > 0.0)) return false;
    const double rjj = sqrt(d);
    R[j * nb + j] = rjj;
    for (int c = j + 1; c < nb; ++c) {
      double s = g[c * nb + j];
      for (int k = 0; k < j; ++k) s -= R[j * nb + k] * R[c * nb + k];
      R[c * nb + j] = s / rjj;
    }
  }
  for (int c = 0; c < nb; ++c) {  // R Ri = I, column c of Ri by back substitution
    for (int i = c; i >= 0; --i) {
      double s = i == c ? 1.0 : 0.0;
      for (int k = i + 1; k <= c; ++k) s -= R[k * nb + i] * Ri[c * nb + k];
      Ri[c * nb + i] = s / R[i * nb + i];
    }
  }
  return true;
}

// out = a b (nb x nb, column-major)
__device__ void mat_mul(int nb, const double* a, const double* b, double* out) {
  double t[kMaxNB * kMaxNB];
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      double s = 0.0;
      for (int k = 0; k < nb; ++k) s += a[k * nb + i] * b[j * nb + k];
      t[j * nb + i] = s;
    }
  for (int i = 0; i < nb * nb; ++i) out[i] = t[i];
}

__device__ bool gauss_inv(int nb, const double* a, double* inv) {
  double m[kMaxNB][2 * kMaxNB];
  for (int r = 0; r < nb; ++r)
    for (int c = 0; c < 2 * nb; ++c) m[r][c] = c < nb ? a[c * nb + r] : (c - nb == r ? 1.0 : 0.0);
  for (int c = 0; c < nb; ++c) {
    int piv = c;
    for (int r = c + 1; r < nb; ++r)
      if (fabs(m[r][c]) > fabs(m[piv][c])) piv = r;
    if (!(m[piv][c] != 0.0)) return false;
    for (int k = 0; k < 2 * nb; ++k) {
      const double t = m[c][k];
      m[c][k] = m[piv][k];
      m[piv][k] = t;
    }
    const double d = m[c][c];
    for (int k = 0; k < 2 * nb; ++k) m[c][k] /= d;
    for (int r = 0; r < nb; ++r) {
      if (r == c) continue;
      const double f = m[r][c];
      if (f != 0.0)
        for (int k = 0; k < 2 * nb; ++k) m[r][k] -= f * m[c][k];
    }
  }
  for (int r = 0; r < nb; ++r)
    for (int c = 0; c < nb; ++c) inv[c * nb + r] = m[r][nb + c];
  return true;
}

enum LzOp : int {
  kOpQr1 = 0,      // Gram -> R1 (kSR1), R1^-1 (kSRinv)
  kOpQr2 = 1,      // Gram -> R2; kSRinv = R2^-1; target = R2 R1 (beta: also beta^T; alpha: transposed)
  kOpInit = 2,     // phibar = beta, rhobar = alpha; resi0 (Right) from beta, resort0 (Right) from alpha
  kOpResi0 = 3,    // resi0 from the Gram diagonal (Left, Central: the unpreconditioned r0)
  kOpGivens = 4,   // block rotation and the coefficient updates
  kOpCheck = 5,    // residual test (Gram diagonals of r and, Right, Ahat^T r)
};

__global__ __launch_bounds__(256) void k_lz_small(int op, int nb, const double* part, int nbp, const double* part2,
                                                  int nbp2, double* sm, KState* ks, int target, int variant,
                                                  double eps, int iter) {
  int* st = ks->st;
  if (st[kLzDone]) return;
  const int K = nb * (nb + 1) / 2;
  __shared__ double gs[2][kGramMax];
  if (part) {
    for (int k = 0; k < K; ++k) {
      const double v = sum_partials(part, nbp, K, k);
      if (threadIdx.x == 0) gs[0][k] = v;
    }
  }
  if (part2) {
    for (int k = 0; k < K; ++k) {
      const double v = sum_partials(part2, nbp2, K, k);
      if (threadIdx.x == 0) gs[1][k] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double g[kMaxNB * kMaxNB];
  for (int c = 0; c < nb; ++c)
    for (int a = 0; a <= c; ++a) g[c * nb + a] = g[a * nb + c] = gs[0][c * (c + 1) / 2 + a];
  switch (op) {
    case kOpQr1: {
      if (!chol_inv(nb, g, sm + kSR1, sm + kSRinv)) { st[kLzBreak] = 1; st[kLzDone] = 1; }
      break;
    }
    case kOpQr2: {
      double R2[kMaxNB * kMaxNB], R[kMaxNB * kMaxNB];
      if (!chol_inv(nb, g, R2, sm + kSRinv)) { st[kLzBreak] = 1; st[kLzDone] = 1; break; }
      mat_mul(nb, R2, sm + kSR1, R);
      double* t = sm + target;
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
          if (target == kSAlpha) t[j * nb + i] = R[i * nb + j];  // alpha = R^T (lanczos.cpp:1117-1124)
          else t[j * nb + i] = R[j * nb + i];
        }
      if (target == kSBeta)
        for (int i = 0; i < nb; ++i)
          for (int j = 0; j < nb; ++j) sm[kSBetaT + j * nb + i] = R[i * nb + j];
      break;
    }
    case kOpInit: {
      for (int i = 0; i < nb * nb; ++i) {
        sm[kSPhibar + i] = sm[kSBeta + i];
        sm[kSRhobar + i] = sm[kSAlpha + i];
      }
      if (variant == 0)
        for (int c = 0; c < nb; ++c) {
          double s = 0.0, so = 0.0;
          for (int i = 0; i < nb; ++i) {
            s += sm[kSBeta + c * nb + i] * sm[kSBeta + c * nb + i];
            so += sm[kSAlpha + c * nb + i] * sm[kSAlpha + c * nb + i];
          }
          sm[kSResi0 + c] = sqrt(s);
          sm[kSResort0 + c] = sqrt(so);
        }
      break;
    }
    case kOpResi0: {
      for (int c = 0; c < nb; ++c) sm[kSResi0 + c] = sqrt(g[c * nb + c]);
      break;
    }
    case kOpGivens: {
      const int n2 = 2 * nb;
      double A[2 * kMaxNB][2 * kMaxNB], Q[2 * kMaxNB][2 * kMaxNB], tau[2 * kMaxNB];
      const double* rhobar = sm + kSRhobar;
      const double* beta = sm + kSBeta;
      for (int i = 0; i < n2; ++i)
        for (int j = 0; j < n2; ++j) A[i][j] = 0.0;
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
          A[i][j] = rhobar[j * nb + i];
          A[nb + i][j] = beta[j * nb + i];
        }
      // Householder QR of the 2nb x 2nb block (dlarfg convention, QrdBlk)
      for (int c = 0; c < n2; ++c) {
        for (int j = 0; j < c; ++j) {
          double s = A[j][c];
          for (int k = j + 1; k < n2; ++k) s += A[k][c] * A[k][j];
          s *= tau[j];
          A[j][c] -= s;
          for (int k = j + 1; k < n2; ++k) A[k][c] -= s * A[k][j];
        }
        double xn2 = 0.0;
        for (int k = c + 1; k < n2; ++k) xn2 += A[k][c] * A[k][c];
        tau[c] = 0.0;
        if (c < n2 - 1 && xn2 != 0.0) {
          const double al = A[c][c];
          const double be = -copysign(sqrt(al * al + xn2), al);
          tau[c] = (be - al) / be;
          const double sc = 1.0 / (al - be);
          for (int k = c + 1; k < n2; ++k) A[k][c] *= sc;
          A[c][c] = be;
        }
      }
      // explicit Q (MvmQBlk of the identity)
      for (int c = 0; c < n2; ++c) {
        double y[2 * kMaxNB];
        for (int i = 0; i < n2; ++i) y[i] = i == c ? 1.0 : 0.0;
        for (int j = n2 - 1; j >= 0; --j) {
          double s = y[j];
          for (int k = j + 1; k < n2; ++k) s += A[k][j] * y[k];
          s = -s * tau[j];
          y[j] += s;
          for (int k = j + 1; k < n2; ++k) y[k] += s * A[k][j];
        }
        for (int i = 0; i < n2; ++i) Q[i][c] = y[i];
      }
      double rho[kMaxNB * kMaxNB], cc[kMaxNB * kMaxNB], ss[kMaxNB * kMaxNB], ccb[kMaxNB * kMaxNB],
          ssb[kMaxNB * kMaxNB];
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) {
          rho[j * nb + i] = i <= j ? A[i][j] : 0.0;
          cc[j * nb + i] = Q[j][i];
          ss[j * nb + i] = Q[nb + j][i];
          ssb[j * nb + i] = -Q[j][nb + i];
          ccb[j * nb + i] = Q[nb + j][nb + i];
        }
      double theta[kMaxNB * kMaxNB], phi[kMaxNB * kMaxNB], t[kMaxNB * kMaxNB], rinv[kMaxNB * kMaxNB];
      mat_mul(nb, ss, sm + kSAlpha, theta);
      mat_mul(nb, ccb, sm + kSAlpha, sm + kSRhobar);
      mat_mul(nb, cc, sm + kSPhibar, phi);
      mat_mul(nb, ssb, sm + kSPhibar, t);
      for (int i = 0; i < nb * nb; ++i) sm[kSPhibar + i] = -t[i];
      if (!gauss_inv(nb, rho, rinv)) { st[kLzBreak] = 1; st[kLzDone] = 1; break; }
      mat_mul(nb, rinv, phi, sm + kSCoef1);
      mat_mul(nb, rinv, theta, sm + kSCoef2);
      break;
    }
    case kOpCheck: {
      int conv = 0;
      for (int c = 0; c < nb; ++c) {
        const double ri = sqrt(g[c * nb + c]);
        sm[kSResi + c] = ri;
        bool ok = ri / sm[kSResi0 + c] < eps;
        if (variant == 0) {
          const int d = c * (c + 1) / 2 + c;
          const double ro = sqrt(gs[1][d]);
          sm[kSResort + c] = ro;
          ok = ok || ro / sm[kSResort0 + c] < eps;
        }
        conv += ok ? 1 : 0;
      }
      st[kLzIter] = iter + 1;
      if (conv == nb) { st[kLzConv] = 1; st[kLzDone] = 1; }
      break;
    }
    default: break;
  }
}

__global__ void k_lz_clear(KState* ks) {
  if (threadIdx.x < 8) ks->st[threadIdx.x] = 0;
}

#define LZ_DISPATCH(nb, call)                          \
  switch (nb) {                                        \
    case 1: { constexpr int NBC = 1; call; } break;    \
    case 2: { constexpr int NBC = 2; call; } break;    \
    case 3: { constexpr int NBC = 3; call; } break;    \
    case 4: { constexpr int NBC = 4; call; } break;    \
    case 5: { constexpr int NBC = 5; call; } break;    \
    case 6: { constexpr int NBC = 6; call; } break;    \
    case 7: { constexpr int NBC = 7; call; } break;    \
    default: { constexpr int NBC = 8; call; } break;   \
  }

struct LzResult {
  int iterations = 0, converged = 0, breakdown = 0;
};

// Block Lanczos over nb (1..8) columns with nonzero initial residuals.
LzResult blanczos_core(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int variant, int nb,
                       const double* b, double* x) {
  const int64_t n = A.d.n;
  const size_t vb = (size_t)n * nb * 8;
  hipStream_t s = ctx.stream;
  const Factor::Image& im = F.to_device(kF64, true);
  double* U = ctx.vec(90, vb);
  double* V = ctx.vec(91, vb);
  double* W = ctx.vec(92, vb);
  double* Q = ctx.vec(93, vb);
  double* S = ctx.vec(94, vb);
  double* t1 = ctx.vec(95, vb);
  double* t2 = ctx.vec(96, vb);
  double* r = ctx.vec(97, vb);
  const int nbr = red_blocks(n);
  double* part = ctx.vec(98, (size_t)nbr * kGramMax * 8);
  double* part2 = ctx.vec(99, (size_t)nbr * kGramMax * 8);
  double* sm = ctx.vec(100, (size_t)kSTotal * 8);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[kLzDone];
  const int ge = red_blocks(n);
  SweepSync* sync = ctx.sweep_sync();

  auto sweep = [&](const DevTri& T, const double* in, double* out, const int* gate) {
    ProfScope ps(ctx, kProfSweep, sweep_bytes(T, nb));
    launch_sweep(T, in, out, kF64, nb, s, gate, sync);
  };
  auto spmv = [&](const DevBsr& M, const double* in, double* out) {
    ProfScope ps(ctx, kProfSpmv, spmv_bytes(M, nb));
    launch_spmv(M, in, out, nullptr, nb, nullptr, nullptr, ctx.rw, s, done);
  };
  // y = Ahat x (t1, t2 scratch; x, y distinct from them)
  auto aop = [&](const double* in, double* out) {
    if (variant == 0) { sweep(im.fwd, in, t1, done); sweep(im.bwd, t1, t2, done); spmv(A.d, t2, out); }
    else if (variant == 1) { spmv(A.d, in, t1); sweep(im.fwd, t1, t2, done); sweep(im.bwd, t2, out, done); }
    else { sweep(im.bwd, in, t1, done); spmv(A.d, t1, t2); sweep(im.fwd, t2, out, done); }
  };
  auto atop = [&](const double* in, double* out) {
    if (variant == 0) { spmv(A.dt, in, t1); sweep(im.fwdt, t1, t2, done); sweep(im.bwdt, t2, out, done); }
    else if (variant == 1) { sweep(im.fwdt, in, t1, done); sweep(im.bwdt, t1, t2, done); spmv(A.dt, t2, out); }
    else { sweep(im.bwdt, in, t1, done); spmv(A.dt, t1, t2); sweep(im.fwdt, t2, out, done); }
  };
  // out = s-map of the iterate: Right U^-1 L^-1 s, Left s, Central U^-1 s.
  // The final map runs ungated (gate null): the stop flag is set by then.
  auto smap = [&](const double* in, double* out, const int* gate) {
    if (variant == 0) { sweep(im.fwd, in, t1, gate); sweep(im.bwd, t1, out, gate); }
    else if (variant == 1) launch_copy((int64_t)n * nb, in, out, s);
    else sweep(im.bwd, in, out, gate);
  };
  auto small = [&](int op, const double* p1, const double* p2, int target, int iter) {
    hipLaunchKernelGGL(k_lz_small, dim3(1), dim3(256), 0, s, op, nb, p1, nbr, p2, nbr, sm, ks, target, variant,
                       opt.eps, iter);
  };
  auto gram = [&](const double* w, double* p) {
    LZ_DISPATCH(nb, hipLaunchKernelGGL(k_lz_gram<NBC>, dim3(nbr), dim3(256), 0, s, n, w, p, done));
  };
  auto mul = [&](const double* X, int coef, const double* base, double sc, double* out) {
    LZ_DISPATCH(nb, hipLaunchKernelGGL(k_lz_mul<NBC>, dim3(ge), dim3(256), 0, s, n, X, sm + coef, base, sc, out,
                                       done));
  };
  // Q R = w (CholeskyQR2): q receives Q, the small state R at `target`.
  auto qr = [&](const double* w, double* q, int target) {
    gram(w, part);
    small(kOpQr1, part, nullptr, 0, 0);
    mul(w, kSRinv, nullptr, 1.0, q);
    gram(q, part);
    small(kOpQr2, part, nullptr, target, 0);
    mul(q, kSRinv, nullptr, 1.0, q);
  };

  hipLaunchKernelGGL(k_lz_clear, dim3(1), dim3(64), 0, s, ks);
  launch_fill((int64_t)n * nb, 0.0, S, s);  // s_0 = 0 (the context's buffers keep earlier solves)
  // r0 = b - A x0, its norms, and the variant's starting block
  launch_spmv(A.d, x, r, b, nb, nullptr, nullptr, ctx.rw, s);
  if (variant != 0) {
    gram(r, part);
    small(kOpResi0, part, nullptr, 0, 0);
    if (variant == 1) { sweep(im.fwd, r, t1, done); sweep(im.bwd, t1, t2, done); }
    else sweep(im.fwd, r, t2, done);
    launch_copy((int64_t)n * nb, t2, r, s);
  }
  qr(r, U, kSBeta);
  atop(U, Q);
  qr(Q, V, kSAlpha);
  small(kOpInit, nullptr, nullptr, 0, 0);
  launch_copy((int64_t)n * nb, V, W, s);
  BCM_HIP(hipGetLastError());
  const int ichk = std::max(1, opt.ichk);
  const int niter = std::max(1, opt.niter);
  LzResult res;
  int k = 0;
  for (k = 0; k < niter; ++k) {
    aop(V, Q);
    mul(U, kSAlpha, Q, -1.0, Q);
    qr(Q, U, kSBeta);
    atop(U, Q);
    mul(V, kSBetaT, Q, -1.0, Q);
    qr(Q, V, kSAlpha);
    small(kOpGivens, nullptr, nullptr, 0, k);
    LZ_DISPATCH(nb, hipLaunchKernelGGL(k_lz_wupd<NBC>, dim3(ge), dim3(256), 0, s, n, W, V, S, sm, done));
    if (k % ichk == 0) {
      smap(S, t2, done);
      launch_axpby((int64_t)n * nb, 1.0, x, 1.0, t2, s);  // t2 = x0 + map(s)
      launch_spmv(A.d, t2, r, b, nb, nullptr, nullptr, ctx.rw, s, done);
      gram(r, part);
      if (variant == 0) {
        atop(r, Q);
        gram(Q, part2);
      }
      small(kOpCheck, part, variant == 0 ? part2 : nullptr, 0, k);
      BCM_HIP(hipGetLastError());
      ctx.pull_state();
      if (ctx.host_state->st[kLzDone]) break;
    }
  }
  ctx.pull_state();
  res.iterations = ctx.host_state->st[kLzBreak] || !ctx.host_state->st[kLzConv] ? std::min(k + 1, niter)
                                                                                  : ctx.host_state->st[kLzIter];
  res.converged = ctx.host_state->st[kLzConv];
  res.breakdown = ctx.host_state->st[kLzBreak];
  // x += map(s) (lanczos.cpp:1593-1596 / 2441 / 3196-3199), also after a breakdown
  smap(S, t2, nullptr);
  launch_axpby((int64_t)n * nb, 1.0, t2, 1.0, x, s);
  return res;
}

}  // namespace

void solve_block_lanczos(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int variant, int nrhs,
                         const double* b, double* x, SolveStats& st, double* resi) {
  if (nrhs < 1 || nrhs > kMaxNB) throw Error(kErrArg, "block Lanczos: 1..8 right-hand sides");
  if (variant < 0 || variant > 2) throw Error(kErrArg, "block Lanczos: variant 0 (Right), 1 (Left), 2 (Central)");
  A.to_device(true);
  if (A.d.n != F.u.n()) throw Error(kErrArg, "matrix and factor sizes differ");
  const int64_t n = A.d.n;
  hipStream_t s = ctx.stream;
  KState* ks = ctx.dstate();
  // columns with r0 = 0 are solved already and stay out of the block
  double* r = ctx.vec(101, (size_t)n * nrhs * 8);
  launch_spmv(A.d, x, r, b, nrhs, nullptr, nullptr, ctx.rw, s);
  {
    const double* xs[kMaxNB];
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
      double* bc = ctx.vec(102, (size_t)n * na * 8);
      double* xc = ctx.vec(103, (size_t)n * na * 8);
      for (int q = 0; q < na; ++q) {
        launch_copy(n, b + (size_t)act[q] * n, bc + (size_t)q * n, s);
        launch_copy(n, x + (size_t)act[q] * n, xc + (size_t)q * n, s);
      }
      bb = bc;
      xx = xc;
    }
    LzResult res = blanczos_core(A, F, ctx, opt, variant, na, bb, xx);
    if (na < nrhs)
      for (int q = 0; q < na; ++q) launch_copy(n, xx + (size_t)q * n, x + (size_t)act[q] * n, s);
    st.iterations = res.iterations;
    st.converged = res.converged;
    st.breakdown = res.breakdown;
  }
  // true residuals of the result
  launch_spmv(A.d, x, r, b, nrhs, nullptr, nullptr, ctx.rw, s);
  {
    const double* xs[2 * kMaxNB];
    for (int c = 0; c < nrhs; ++c) {
      xs[c] = r + (size_t)c * n;
      xs[nrhs + c] = b + (size_t)c * n;
    }
    launch_multidot((int)n, nrhs, xs, xs, &ks->sc[16], ctx.rw, s);
    launch_multidot((int)n, nrhs, xs + nrhs, xs + nrhs, &ks->sc[24], ctx.rw, s);
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
