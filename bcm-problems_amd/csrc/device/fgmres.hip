// Flexible GMRES(m) on the GPU with a mixed-precision preconditioner.
//
// The outer method is right-preconditioned flexible GMRES (Saad 1993): the
// preconditioned directions Z_j = M^-1 V_j are stored, so M^-1 may change
// from step to step -- here it is the IC2/ILU2 sweep pair applied in FP64,
// FP32 or BF16 storage (FP32 accumulation), which makes the outer loop an
// iterative refinement of the low-precision preconditioner. Everything the
// outer method owns stays FP64 on the device: the Krylov basis, the
// Gram-Schmidt dots and updates (classical Gram-Schmidt with one
// re-orthogonalisation pass, CGS2), the Hessenberg column, the Givens
// rotations, the least-squares solve, the solution update and the
// true-residual recompute at every restart.
//
// Control stays on the device: one-block kernels finalise the reductions and
// run the Hessenberg/Givens step; once the residual estimate meets the
// tolerance (or the cycle is full) they raise a cycle-end flag and every
// step kernel still queued returns at entry. The host polls the flag every
// few steps, and once per cycle.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

constexpr int kMaxRestart = 64;
constexpr int kGsBlocks = 512;  // blocks of the Gram-Schmidt passes (partials per value)

// KState slots used by FGMRES
enum : int {
  kSlotB2 = 20, kSlotR02 = 21, kSlotBeta = 24,
};
enum : int {
  kStDone = 0, kStSteps = 1, kStBreak = 2, kStConv = 4, kStCycleEnd = 5, kStCycleLen = 6,
};

// Offsets in the FGMRES scalar workspace (doubles), for restart length m.
struct GmLayout {
  int m;
  __host__ __device__ int H() const { return 0; }                      // (m+1) x m, column j at H + j*(m+1)
  __host__ __device__ int cs() const { return (m + 1) * m; }
  __host__ __device__ int sn() const { return cs() + m; }
  __host__ __device__ int g() const { return sn() + m; }               // m+1
  __host__ __device__ int y() const { return g() + m + 1; }            // m
  __host__ __device__ int h1() const { return y() + m; }               // first Gram-Schmidt pass, m+1
  __host__ __device__ int h2() const { return h1() + m + 1; }          // second pass, m+1
  __host__ __device__ int inv() const { return h2() + m + 1; }         // 1 / (norm of the new vector)
  __host__ __device__ int size() const { return inv() + 1; }
};

struct DotPairs {
  const double* x[4];
  const double* y[4];
};

__global__ void k_state_reset_all(KState* ks) {
  const int t = threadIdx.x;
  if (t < 32) ks->sc[t] = 0.0;
  if (t < 8) ks->st[t] = 0;
}

// sc[slot0 + k] = sum of the partials of value k (one block per value).
__global__ __launch_bounds__(256) void k_fin_to_state(KState* ks, const double* partials, int nb, int K,
                                                      int slot0) {
  const double s = sum_partials(partials, nb, K, blockIdx.x);
  if (threadIdx.x == 0) ks->sc[slot0 + blockIdx.x] = s;
}

__device__ double gm_threshold(const KState* ks, int ittype, double eps) {
  if (ittype == 0) return eps * sqrt(ks->sc[kSlotR02]);
  if (ittype == 1) return eps;
  return eps * sqrt(ks->sc[kSlotB2]);
}

// Per-block partial dots of K pairs (x_k, y_k).
template <int K>
__global__ __launch_bounds__(256) void k_pdots(int64_t n, DotPairs args, double* partials) {
  double v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += args.x[k][i] * args.y[k][i];
  }
  block_partials<K>(v, partials);
}

// Cycle start: partials hold (r.r, b.b). Sets beta = ||r||, g = beta e1 and
// decides whether the current x already meets the tolerance.
__global__ __launch_bounds__(256) void k_gm_start(KState* ks, double* gm, GmLayout L, const double* partials,
                                                  int nb, int first, double eps, int ittype, int steps_left) {
  const double rr = sum_partials(partials, nb, 2, 0);
  const double bb = sum_partials(partials, nb, 2, 1);
  if (threadIdx.x != 0) return;
  double* sc = ks->sc;
  int* st = ks->st;
  if (first) {
    sc[kSlotB2] = bb;
    sc[kSlotR02] = rr;
  }
  const double beta = sqrt(rr);
  sc[kSlotBeta] = beta;
  st[kStCycleEnd] = 0;
  st[kStCycleLen] = 0;
  st[kStConv] = 0;
  for (int i = 0; i <= L.m; ++i) gm[L.g() + i] = 0.0;
  gm[L.g()] = beta;
  gm[L.inv()] = beta > 0.0 ? 1.0 / beta : 0.0;
  if (!(beta > gm_threshold(ks, ittype, eps))) {
    st[kStConv] = 1;
    st[kStDone] = 1;
    st[kStCycleEnd] = 1;
  } else if (steps_left <= 0) {
    st[kStDone] = 1;
    st[kStCycleEnd] = 1;
  }
}

// dst = src * gm[inv] (the new basis vector), unless the cycle has ended.
__global__ void k_gm_scale(int64_t n, const KState* ks, const double* gm, int inv_off, const double* src,
                           double* dst) {
  if (ks->st[kStCycleEnd]) return;
  const double s = gm[inv_off];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i] * s;
}

// First Gram-Schmidt pass: partial dots V_k . w for k < nv.
template <int NV>
__global__ __launch_bounds__(256) void k_gs_dot(int64_t n, const KState* ks, const double* __restrict__ V, int nv,
                                                const double* __restrict__ w, double* partials) {
  if (ks->st[kStCycleEnd]) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double wi = w[i];
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k < nv) acc[k] += V[(size_t)k * n + i] * wi;
  }
  block_partials<NV>(acc, partials);
}

// h[k] = sum of the per-block partials of value k (one block per value).
__global__ __launch_bounds__(256) void k_gs_fin(const KState* ks, const double* partials, int nb, int stride_k,
                                                double* h) {
  if (ks->st[kStCycleEnd]) return;
  const double s = sum_partials(partials, nb, stride_k, blockIdx.x);
  if (threadIdx.x == 0) h[blockIdx.x] = s;
}

// w -= sum_k h[k] V_k (k ascending); then either the partial dots V_k . w
// of the re-orthogonalisation pass (NORM false) or the partial ||w||^2.
template <int NV, bool NORM>
__global__ __launch_bounds__(256) void k_gs_update(int64_t n, const KState* ks, const double* __restrict__ V, int nv,
                                                   const double* __restrict__ h, double* __restrict__ w,
                                                   double* partials) {
  if (ks->st[kStCycleEnd]) return;
  constexpr int NA = NORM ? 1 : NV;
  double hk[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) hk[k] = k < nv ? h[k] : 0.0;
  double acc[NA];
#pragma unroll
  for (int k = 0; k < NA; ++k) acc[k] = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double vk[NV];
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      vk[k] = k < nv ? V[(size_t)k * n + i] : 0.0;
      if (k < nv) wi -= hk[k] * vk[k];
    }
    w[i] = wi;
    if constexpr (NORM) {
      acc[0] += wi * wi;
    } else {
#pragma unroll
      for (int k = 0; k < NV; ++k)
        if (k < nv) acc[k] += vk[k] * wi;
    }
  }
  block_partials<NA>(acc, partials);
}

// Arnoldi step j: partials hold ||w||^2 of the orthogonalised vector. Builds
// column j of H (h1 + h2, norm), applies the earlier rotations, forms the new
// one, updates g and decides whether the cycle ends.
__global__ __launch_bounds__(256) void k_gm_step(KState* ks, double* gm, GmLayout L, const double* partials,
                                                 int nb, int j, double eps, int ittype, int steps_left) {
  if (ks->st[kStCycleEnd]) return;
  const double ww = sum_partials(partials, nb, 1, 0);
  if (threadIdx.x != 0) return;
  int* st = ks->st;
  const int m = L.m;
  double* col = gm + L.H() + (size_t)j * (m + 1);
  const double* h1 = gm + L.h1();
  const double* h2 = gm + L.h2();
  for (int k = 0; k <= j; ++k) col[k] = h1[k] + h2[k];
  const double hn = sqrt(ww);
  col[j + 1] = hn;
  double* cs = gm + L.cs();
  double* sn = gm + L.sn();
  double* g = gm + L.g();
  for (int i = 0; i < j; ++i) {
    const double a = col[i], b = col[i + 1];
    col[i] = cs[i] * a + sn[i] * b;
    col[i + 1] = -sn[i] * a + cs[i] * b;
  }
  const double a = col[j], b = col[j + 1];
  const double rr = hypot(a, b);
  if (!(rr > 0.0) || !isfinite(rr)) {  // A Z_j lies in the span already and is zero: breakdown
    st[kStBreak] = 1;
    st[kStCycleEnd] = 1;
    st[kStDone] = 1;
    return;
  }
  cs[j] = a / rr;
  sn[j] = b / rr;
  col[j] = rr;
  col[j + 1] = 0.0;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
  st[kStSteps] += 1;
  st[kStCycleLen] = j + 1;
  gm[L.inv()] = hn > 0.0 ? 1.0 / hn : 0.0;
  if (!(fabs(g[j + 1]) > gm_threshold(ks, ittype, eps))) {
    st[kStConv] = 1;
    st[kStCycleEnd] = 1;
  } else if (j + 1 >= m || hn == 0.0) {
    st[kStCycleEnd] = 1;  // full cycle, or the Krylov space is invariant (estimate is exact)
  }
  if (j + 1 >= steps_left) {  // iteration budget spent
    st[kStCycleEnd] = 1;
    st[kStDone] = 1;
  }
}

// Least squares: y = R^-1 g over the cycle's columns (upper triangular R in H).
__global__ void k_gm_solve(const KState* ks, double* gm, GmLayout L) {
  const int jn = ks->st[kStCycleLen];
  const int m = L.m;
  const double* H = gm + L.H();
  const double* g = gm + L.g();
  double* y = gm + L.y();
  for (int i = jn - 1; i >= 0; --i) {
    double s = g[i];
    for (int k = i + 1; k < jn; ++k) s -= H[(size_t)k * (m + 1) + i] * y[k];
    y[i] = s / H[(size_t)i * (m + 1) + i];
  }
}

// x += sum_{k < cycle length} y[k] Z_k (k ascending).
__global__ __launch_bounds__(256) void k_gm_update_x(int64_t n, const KState* ks, const double* gm, GmLayout L,
                                                     const double* __restrict__ Z, double* __restrict__ x) {
  const int jn = ks->st[kStCycleLen];
  if (jn == 0) return;
  const double* y = gm + L.y();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double xi = x[i];
    for (int k = 0; k < jn; ++k) xi += y[k] * Z[(size_t)k * n + i];
    x[i] = xi;
  }
}

int grid_elems(int64_t n, int cap) {
  const int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

template <int NV>
void gs_pass(int64_t n, KState* ks, const double* V, int nv, double* w, double* gm, const GmLayout& L,
             double* partials, int nb, hipStream_t s) {
  hipLaunchKernelGGL(k_gs_dot<NV>, dim3(nb), dim3(256), 0, s, n, ks, V, nv, w, partials);
  hipLaunchKernelGGL(k_gs_fin, dim3(nv), dim3(256), 0, s, ks, partials, nb, NV, gm + L.h1());
  hipLaunchKernelGGL((k_gs_update<NV, false>), dim3(nb), dim3(256), 0, s, n, ks, V, nv, gm + L.h1(), w, partials);
  hipLaunchKernelGGL(k_gs_fin, dim3(nv), dim3(256), 0, s, ks, partials, nb, NV, gm + L.h2());
  hipLaunchKernelGGL((k_gs_update<NV, true>), dim3(nb), dim3(256), 0, s, n, ks, V, nv, gm + L.h2(), w, partials);
}

// Orthogonalises w against V_0..V_{nv-1} (CGS2); leaves the partial ||w||^2.
void gram_schmidt(int64_t n, KState* ks, const double* V, int nv, double* w, double* gm, const GmLayout& L,
                  double* partials, int nb, hipStream_t s) {
  if (nv <= 8) gs_pass<8>(n, ks, V, nv, w, gm, L, partials, nb, s);
  else if (nv <= 16) gs_pass<16>(n, ks, V, nv, w, gm, L, partials, nb, s);
  else if (nv <= 32) gs_pass<32>(n, ks, V, nv, w, gm, L, partials, nb, s);
  else gs_pass<64>(n, ks, V, nv, w, gm, L, partials, nb, s);
}

}  // namespace

void solve_fgmres(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
                  SolveStats& st) {
  const int m = opt.restart;
  if (m < 1 || m > kMaxRestart) throw Error(kErrArg, "FGMRES restart length must be in 1..64");
  const int64_t n = A.d.n;
  const size_t vb = (size_t)n * 8;
  GmLayout L{m};
  double* V = ctx.vec(30, vb * (m + 1));
  double* Z = ctx.vec(31, vb * m);
  double* r = ctx.vec(32, vb);
  double* gm = ctx.vec(33, (size_t)L.size() * 8);
  double* partials = ctx.vec(34, (size_t)kGsBlocks * 64 * 8);
  KState* ks = ctx.dstate();
  hipStream_t s = ctx.stream;
  const int nb = grid_elems(n, kGsBlocks);
  const int ge = grid_elems(n, 1024);
  const int* cycle_end = &ks->st[kStCycleEnd];
  const int poll = std::max(1, opt.check_every);
  hipLaunchKernelGGL(k_state_reset_all, dim3(1), dim3(64), 0, s, ks);
  int steps = 0, cycles = 0;
  bool first = true;
  while (true) {
    // r = b - A x; (r.r, b.b); beta, g = beta e1, V_0 = r / beta
    launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);
    DotPairs dp{};
    dp.x[0] = r; dp.y[0] = r; dp.x[1] = b; dp.y[1] = b;
    hipLaunchKernelGGL(k_pdots<2>, dim3(nb), dim3(256), 0, s, n, dp, partials);
    hipLaunchKernelGGL(k_gm_start, dim3(1), dim3(256), 0, s, ks, gm, L, partials, nb, first ? 1 : 0, opt.eps,
                       opt.ittype, opt.niter - steps);
    hipLaunchKernelGGL(k_gm_scale, dim3(ge), dim3(256), 0, s, n, ks, gm, L.inv(), r, V);
    first = false;
    for (int j = 0; j < m; ++j) {
      double* vj = V + (size_t)j * n;
      double* zj = Z + (size_t)j * n;
      double* w = V + (size_t)(j + 1) * n;
      precond_apply(F, ctx, vj, zj, opt.prec, false, 1, cycle_end);
      {
        ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
        launch_spmv(A.d, zj, w, nullptr, 1, nullptr, nullptr, ctx.rw, s, cycle_end);
      }
      gram_schmidt(n, ks, V, j + 1, w, gm, L, partials, nb, s);
      hipLaunchKernelGGL(k_gm_step, dim3(1), dim3(256), 0, s, ks, gm, L, partials, nb, j, opt.eps, opt.ittype,
                         opt.niter - steps);
      hipLaunchKernelGGL(k_gm_scale, dim3(ge), dim3(256), 0, s, n, ks, gm, L.inv(), w, w);
      if ((j + 1) % poll == 0 || j + 1 == m) {
        ctx.pull_state();
        if (ctx.host_state->st[kStCycleEnd]) break;
      }
    }
    hipLaunchKernelGGL(k_gm_solve, dim3(1), dim3(1), 0, s, ks, gm, L);
    hipLaunchKernelGGL(k_gm_update_x, dim3(ge), dim3(256), 0, s, n, ks, gm, L, Z, x);
    ctx.pull_state();
    const KState& hs = *ctx.host_state;
    steps = hs.st[kStSteps];
    if (hs.st[kStCycleLen] > 0) ++cycles;
    if (hs.st[kStDone] || hs.st[kStBreak] || hs.st[kStCycleLen] == 0) break;
  }
  st.restarts = std::max(0, cycles - 1);
  const KState& hs = *ctx.host_state;
  st.iterations = hs.st[kStSteps];
  st.breakdown = hs.st[kStBreak];
  // true residual of the final iterate
  launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);
  DotPairs dp{};
  dp.x[0] = r; dp.y[0] = r; dp.x[1] = b; dp.y[1] = b;
  hipLaunchKernelGGL(k_pdots<2>, dim3(nb), dim3(256), 0, s, n, dp, partials);
  hipLaunchKernelGGL(k_fin_to_state, dim3(2), dim3(256), 0, s, ks, partials, nb, 2, 22);
  ctx.pull_state();
  const double rr = ctx.host_state->sc[22], bb = ctx.host_state->sc[23];
  st.relres_true = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
  st.relres_rec = st.relres_true;
  const double thr = opt.ittype == 0 ? opt.eps * std::sqrt(ctx.host_state->sc[kSlotR02])
                   : opt.ittype == 1 ? opt.eps : opt.eps * std::sqrt(bb);
  st.converged = std::sqrt(rr) <= thr ? 1 : 0;
}

}  // namespace bcm
