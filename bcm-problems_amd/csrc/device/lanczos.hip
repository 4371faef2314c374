// Two-sided Lanczos (BiCG-type) solver of the reference, serial form
// (CSVector::Lanczos, lanczos.cpp:34-371), on the GPU.
//
// Per iteration, as the reference: Img = U^-1 L^-1 A delta (SolveL with L,
// SolveU with U), the scalars delDel = (delta, Delta), gamma = delDel /
// (Img, Delta) and the three-term coefficient ro; the solution update; the
// true residual b - A x with its max-norm and 2-norm, keeping the iterate of
// smallest max-norm residual (the reference's goodsol); then the two
// direction updates, the second one through the transposed preconditioner
// (U.SolveL then L.SolveU) and A^T. All scalars stay on the device; the
// stopping test and breakdown check run in one-block kernels and every kernel
// queued past the end returns at entry.
#include <algorithm>
#include <cmath>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

enum : int {  // KState::sc slots
  kLzDelDel = 0, kLzDelDel1 = 1, kLzGamma = 2, kLzGamma1 = 3, kLzRo = 4, kLzRoGamma = 5, kLzErrMin = 6,
  kLzResIni = 7, kLzDot0 = 8, kLzDot1 = 9, kLzErr = 10, kLzRes2 = 11,
};
enum : int { kLzDone = 0, kLzIter = 1, kLzBreak = 2, kLzGood = 5 };

// Per-block max |r_i| and sum r_i^2 (max-norm residual with the unit norm
// vector, MaxProdSVect svector.cpp:376-385, and the 2-norm).
__global__ __launch_bounds__(256) void k_lz_norms(int64_t n, const KState* ks, const double* __restrict__ r,
                                                  double* partials) {
  if (ks->st[kLzDone]) return;
  double mx = 0.0, ss = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = r[i];
    mx = fmax(mx, fabs(v));
    ss += v * v;
  }
  __shared__ double smx[16];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) smx[wave] = mx;
  const double v2[1] = {ss};
  block_partials<1>(v2, partials + gridDim.x);  // sums after the maxima (block_partials synchronises the block)
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, smx[w]);
    partials[blockIdx.x] = m;
  }
}

// Scalar step after (delta.Delta, Img.Delta): gamma, ro and the update coefficients.
__global__ __launch_bounds__(256) void k_lz_gamma(KState* ks, const double* partials, int nb, int k) {
  if (ks->st[kLzDone]) return;
  const double dd = sum_partials(partials, nb, 2, 0);
  const double dt = sum_partials(partials, nb, 2, 1);
  if (threadIdx.x != 0) return;
  double* sc = ks->sc;
  sc[kLzGamma1] = sc[kLzGamma];
  sc[kLzDelDel1] = sc[kLzDelDel];
  sc[kLzDelDel] = dd;
  sc[kLzGamma] = dd / dt;
  if (k > 1) sc[kLzRo] = 1.0 / (1.0 - (sc[kLzGamma] / sc[kLzGamma1]) * (sc[kLzDelDel] / sc[kLzDelDel1]) / sc[kLzRo]);
  sc[kLzRoGamma] = sc[kLzRo] * sc[kLzGamma];
}

// x1 <- x; x <- ro x + (1 - ro) x1_old + rogamma delta.
__global__ void k_lz_sol(int64_t n, const KState* ks, double* __restrict__ x, double* __restrict__ x1,
                         const double* __restrict__ delta) {
  if (ks->st[kLzDone]) return;
  const double ro = ks->sc[kLzRo], rg = ks->sc[kLzRoGamma], un = 1. - ro;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double xi = x[i];
    const double p = xi * ro + x1[i] * un;
    x1[i] = xi;
    x[i] = p + delta[i] * rg;
  }
}

// Residual norms of the new iterate, goodsol bookkeeping and the stopping test.
__global__ __launch_bounds__(256) void k_lz_check(KState* ks, const double* partials, int nb, int k, int niter,
                                                  double eps, int sttype) {
  if (ks->st[kLzDone]) return;
  __shared__ double smx[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, partials[i]);
  smx[threadIdx.x] = m;
  __syncthreads();
  const double res2 = sum_partials(partials + nb, nb, 1, 0);
  if (threadIdx.x != 0) return;
  double errloc = 0.0;
  for (int i = 0; i < (int)blockDim.x; ++i) errloc = fmax(errloc, smx[i]);
  double* sc = ks->sc;
  int* st = ks->st;
  const double resi = sqrt(res2);
  st[kLzGood] = 0;
  if (errloc < sc[kLzErrMin]) {
    st[kLzGood] = 1;
    sc[kLzErrMin] = errloc;
  }
  st[kLzIter] = k;
  sc[kLzErr] = errloc;
  sc[kLzRes2] = resi;
  const bool brk = fabs(sc[kLzDelDel]) < 1.e-100 && fabs(sc[kLzDelDel1]) < 1.e-100;
  bool met;
  if (sttype == 2) met = sc[kLzErrMin] <= eps;
  else if (sttype == 0) met = resi <= eps;
  else met = resi <= eps * sc[kLzResIni];
  if (met || brk || k >= niter) {
    st[kLzDone] = 1;
    st[kLzBreak] = brk ? 1 : 0;
    st[4] = met ? 1 : 0;
  }
}

__global__ void k_lz_keep(int64_t n, const KState* ks, const double* __restrict__ x, double* __restrict__ good) {
  if (!ks->st[kLzGood]) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) good[i] = x[i];
}

// v1 <- v; v <- ro v + (1 - ro) v1_old - rogamma img.
__global__ void k_lz_dir(int64_t n, const KState* ks, double* __restrict__ v, double* __restrict__ v1,
                         const double* __restrict__ img) {
  if (ks->st[kLzDone]) return;
  const double ro = ks->sc[kLzRo], rg = ks->sc[kLzRoGamma], un = 1. - ro;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double vi = v[i];
    const double p = vi * ro + v1[i] * un;
    v1[i] = vi;
    v[i] = p - img[i] * rg;
  }
}

// Initial state: errmin = max|r0|, resini = ||r0||, ro = 1, gamma = delDel = 0.
__global__ __launch_bounds__(256) void k_lz_init(KState* ks, const double* partials, int nb) {
  __shared__ double smx[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, partials[i]);
  smx[threadIdx.x] = m;
  __syncthreads();
  const double res2 = sum_partials(partials + nb, nb, 1, 0);
  if (threadIdx.x != 0) return;
  double e = 0.0;
  for (int i = 0; i < (int)blockDim.x; ++i) e = fmax(e, smx[i]);
  for (int i = 0; i < 32; ++i) ks->sc[i] = 0.0;
  for (int i = 0; i < 8; ++i) ks->st[i] = 0;
  ks->sc[kLzErrMin] = e;
  ks->sc[kLzResIni] = sqrt(res2);
  ks->sc[kLzRo] = 1.0;
}

__global__ void k_lz_clear(KState* ks) {
  if (threadIdx.x < 8) ks->st[threadIdx.x] = 0;
}

int grid_n(int64_t n, int cap) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)); }

}  // namespace

void solve_lanczos(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
                   SolveStats& st) {
  if (F.kind != 1) throw Error(kErrArg, "Lanczos needs ILU2 factors (L and U)");
  A.to_device(true);
  const int64_t n = A.d.n;
  const size_t vb = (size_t)n * 8;
  double* r = ctx.vec(50, vb);
  double* good = ctx.vec(51, vb);
  double* x1 = ctx.vec(52, vb);
  double* Dl = ctx.vec(53, vb);   // Delta
  double* Dl1 = ctx.vec(54, vb);  // Delta1
  double* dl = ctx.vec(55, vb);   // delta
  double* dl1 = ctx.vec(56, vb);  // delta1
  double* img = ctx.vec(57, vb);
  double* tmp = ctx.vec(58, vb);
  const int nb = grid_n(n, 1024);
  double* partials = ctx.vec(59, (size_t)nb * 2 * 8 + 64);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[kLzDone];
  hipStream_t s = ctx.stream;
  const int ge = grid_n(n, 1024);
  // sttype of the reference: 0 absolute, 1 relative to ||r0||, 2 max-norm; 3 (extension) relative to ||b||
  const int sttype = opt.ittype == 0 ? 1 : opt.ittype == 2 ? 2 : opt.ittype == 3 ? 3 : 0;
  double eps = opt.eps;
  // r = b - A x, norms, goodsol = x; Delta = r; delta = M^-1 r (the flags of a previous solve are cleared first)
  hipLaunchKernelGGL(k_lz_clear, dim3(1), dim3(64), 0, s, ks);
  launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);
  hipLaunchKernelGGL(k_lz_norms, dim3(nb), dim3(256), 0, s, n, ks, r, partials);
  hipLaunchKernelGGL(k_lz_init, dim3(1), dim3(256), 0, s, ks, partials, nb);
  if (sttype == 3) {  // ||r|| <= eps ||b||: an absolute test against eps * ||b||
    const double* xs[1] = {b};
    launch_multidot((int)n, 1, xs, xs, &ks->sc[12], ctx.rw, s);
    ctx.pull_state();
    eps *= std::sqrt(ctx.host_state->sc[12]);
  }
  launch_copy(n, x, good, s);
  launch_copy(n, r, Dl, s);
  launch_fill(n, 0.0, x1, s);
  launch_fill(n, 0.0, Dl1, s);
  launch_fill(n, 0.0, dl1, s);
  precond_apply(F, ctx, r, dl, kF64, false, 1);
  const int st_sttype = sttype == 3 ? 0 : sttype;
  const bool exact = opt.check_every == 1;
  int k = 1;
  const int kmax = std::max(opt.niter, 2);  // the reference's loop: k = 1 .. niter-1
  while (k < kmax) {
    for (int c = 0; c < std::max(1, opt.check_every) && k < kmax; ++c, ++k) {
      {
        ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
        launch_spmv(A.d, dl, r, nullptr, 1, nullptr, nullptr, ctx.rw, s, done);
      }
      precond_apply(F, ctx, r, img, kF64, false, 1, done);
      {
        const double* xs[2] = {dl, img};
        const double* ys[2] = {Dl, Dl};
        launch_multidot((int)n, 2, xs, ys, kDeferred, ctx.rw, s, done);
        hipLaunchKernelGGL(k_lz_gamma, dim3(1), dim3(256), 0, s, ks, ctx.rw.partials.as<double>(),
                           ctx.rw.pending_blocks, k);
        ctx.rw.pending_k = 0;
      }
      hipLaunchKernelGGL(k_lz_sol, dim3(ge), dim3(256), 0, s, n, ks, x, x1, dl);
      {
        ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
        launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s, done);
      }
      hipLaunchKernelGGL(k_lz_norms, dim3(nb), dim3(256), 0, s, n, ks, r, partials);
      hipLaunchKernelGGL(k_lz_check, dim3(1), dim3(256), 0, s, ks, partials, nb, k, opt.niter, eps, st_sttype);
      hipLaunchKernelGGL(k_lz_keep, dim3(ge), dim3(256), 0, s, n, ks, x, good);
      hipLaunchKernelGGL(k_lz_dir, dim3(ge), dim3(256), 0, s, n, ks, dl, dl1, img);
      precond_apply(F, ctx, Dl, tmp, kF64, true, 1, done);  // U.SolveL then L.SolveU
      {
        ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.dt, 1));
        launch_spmv(A.dt, tmp, img, nullptr, 1, nullptr, nullptr, ctx.rw, s, done);
      }
      hipLaunchKernelGGL(k_lz_dir, dim3(ge), dim3(256), 0, s, n, ks, Dl, Dl1, img);
      if (exact) {
        ctx.pull_state();
        if (ctx.host_state->st[kLzDone]) break;
      }
    }
    ctx.pull_state();
    if (ctx.host_state->st[kLzDone]) break;
  }
  ctx.pull_state();
  const KState& hs = *ctx.host_state;
  st.iterations = hs.st[kLzIter];
  st.breakdown = hs.st[kLzBreak];
  launch_copy(n, good, x, s);
  // true residual of the returned (good) iterate
  launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);
  const double* xs[2] = {r, b};
  launch_multidot((int)n, 2, xs, xs, &ks->sc[22], ctx.rw, s);
  ctx.pull_state();
  const double rr = ctx.host_state->sc[22], bb = ctx.host_state->sc[23];
  st.relres_true = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  st.relres_rec = hs.sc[kLzRes2] / (bb > 0 ? std::sqrt(bb) : 1.0);
  st.converged = hs.st[4];  // the reference's stopping test was met (the good iterate is returned)
}

}  // namespace bcm
