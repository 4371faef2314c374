// GPU-resident Krylov drivers: preconditioned CG (the reference's serial Cg,
// cg.cpp:31-349, and its parallel form cg.cpp:355-771) and right-preconditioned
// BiCGStab; plus the device images of matrices and factors and the
// preconditioner application.
//
// Every recurrence scalar lives in device memory (KState). A solver iteration
// is a fixed sequence of launches; scalar updates run in one-thread kernels
// between the vector kernels, so no vector kernel ever races with a scalar
// write. When the device-side test marks the solve finished, every kernel
// still queued returns at entry, so the host only polls the state every few
// iterations and the reported iteration count is exact.
#include <chrono>
#include <cmath>
#include <cstring>

#include "dev_common.hpp"
#include "solver.hpp"

namespace bcm {

namespace {

enum ScalarOp : int {
  kOpPcgCheck = 1,
  kOpPcgAlpha = 2,
  kOpBicgInit = 10,
  kOpBicgAlpha = 11,
  kOpBicgSCheck = 12,
  kOpBicgOmega = 13,
  kOpBicgNext = 14,
  kOpCgParAlpha = 20,
  kOpCgParCheck = 21,
  kOpCgParBeta = 22,
};

// Scalar slots (KState::sc)
enum : int {
  // PCG
  kGamold = 1, kPq = 3, kEnrsq0 = 4, kBeta = 5, kAlpha = 6, kEnrsq = 8, kGamma = 9,
  // BiCGStab
  kRho = 0, kRhoOld = 1, kBAlpha = 2, kOmega = 3, kRv = 4, kSs = 5, kTs = 6, kTt = 7, kRr = 8,
  kRhoNew = 10,
  // parallel-form CG
  kCpUau = 11, kCpRu = 12, kCpAcoef = 13, kCpBcoef = 14, kCpRr = 15, kCpVau = 16, kCpThr2 = 17,
  // shared
  kBnorm2 = 20, kR0norm2 = 21, kTrue0 = 22, kTrue1 = 23,
  // PCG, ittype 2: max_i |r_i| of the residual at the check (MaxProdSVect with a unit norm vector)
  kErrloc = 24,
};

__device__ double threshold(const KState* ks, int ittype, double eps2) {
  if (ittype == 0) return eps2 * ks->sc[kR0norm2];
  if (ittype == 1) return eps2;
  return eps2 * ks->sc[kBnorm2];
}

__global__ void k_state_reset(KState* ks, int keep_scalars) {
  const int t = threadIdx.x;
  if (!keep_scalars && t < 32) ks->sc[t] = 0.0;
  if (t < 8) ks->st[t] = 0;
}

// One-block scalar step. It first adds the pending per-block partials of the
// preceding reduction (K values, stored from slot0 on), then thread 0 runs
// the recurrence update `op`.
__global__ __launch_bounds__(256) void k_scalar(int op, KState* ks, double eps2, int ittype, int niter,
                                                const double* partials, int nb, int K, int slot0, int stamp) {
  double* sc = ks->sc;
  int* st = ks->st;
  if (st[0]) return;
  double v[kMaxDots];
#pragma unroll
  for (int k = 0; k < kMaxDots; ++k)
    if (k < K) v[k] = sum_partials(partials, nb, K, k);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < K && k < kMaxDots; ++k) sc[slot0 + k] = v[k];
  switch (op) {
    case kOpPcgCheck: {
      const int k = ++st[1];
      if (k == 1) sc[kEnrsq0] = sc[kEnrsq];
      if (ittype == 2) {  // cg.cpp:241: errmin <= |eps|, eps2 carries |eps|
        if (!(sc[kErrloc] > eps2)) { st[0] = 1; st[4] = 1; return; }
      } else {
        const double thr = ittype == 0 ? eps2 * sc[kEnrsq0] : threshold(ks, ittype, eps2);
        if (!(sc[kEnrsq] > thr)) { st[0] = 1; st[4] = 1; return; }
      }
      if (k >= niter) { st[0] = 1; return; }
      if (!(sc[kGamma] > 0.0)) { st[0] = 1; st[2] = 1; return; }
      sc[kBeta] = k > 1 ? sc[kGamma] / sc[kGamold] : 0.0;
      break;
    }
    case kOpPcgAlpha: {
      const double alpha = sc[kGamma] / sc[kPq];
      if (!(alpha >= 0.0)) { st[0] = 1; st[2] = 1; return; }  // "Matrix is not positive definite"
      sc[kAlpha] = alpha;
      sc[kGamold] = sc[kGamma];
      break;
    }
    case kOpBicgInit: {  // slots 20..22 hold b.b, r0.r0, rhat.r0
      sc[kRr] = sc[kR0norm2];
      sc[kRho] = sc[kTrue0];
      sc[kRhoOld] = 1.0;
      sc[kBAlpha] = 1.0;
      sc[kOmega] = 1.0;
      if (!(sc[kRr] > threshold(ks, ittype, eps2))) { st[0] = 1; st[4] = 1; }
      break;
    }
    case kOpBicgAlpha: {
      const double a = sc[kRho] / sc[kRv];
      if (!(std::isfinite(a)) || sc[kRv] == 0.0) { st[0] = 1; st[2] = 1; return; }
      sc[kBAlpha] = a;
      break;
    }
    case kOpBicgSCheck: {  // converged on s: the x update by alpha alone (k_bicg_xs with this stamp) ends it
      if (!(sc[kSs] > threshold(ks, ittype, eps2))) { st[0] = 1; st[3] = stamp; st[4] = 1; ++st[1]; }
      break;
    }
    case kOpBicgOmega: {
      const double w = sc[kTs] / sc[kTt];
      if (!(std::isfinite(w)) || w == 0.0) { st[0] = 1; st[2] = 1; return; }
      sc[kOmega] = w;
      break;
    }
    case kOpBicgNext: {  // slots kRr, kRr + 1 hold r.r and rhat.r
      const int k = ++st[1];
      sc[kRhoOld] = sc[kRho];
      sc[kRho] = sc[kRr + 1];
      if (!(sc[kRr] > threshold(ks, ittype, eps2))) { st[0] = 1; st[4] = 1; return; }
      if (k >= niter) { st[0] = 1; return; }
      if (sc[kRho] == 0.0 || !std::isfinite(sc[kRho])) { st[0] = 1; st[2] = 1; return; }
      break;
    }
    case kOpCgParAlpha: {  // slots kCpUau, kCpRu hold (Au, u), (r, u)
      if (sc[kCpUau] < 0.0) { st[0] = 1; st[2] = 1; return; }  // "Matrix is not positive definite"
      sc[kCpAcoef] = (sc[kCpRu] * -1.0) / sc[kCpUau];
      ++st[1];
      break;
    }
    case kOpCgParCheck: {  // slot kCpRr holds (r, r) after the residual update
      if (!(sc[kCpRr] > sc[kCpThr2])) { st[0] = 1; st[4] = 1; st[6] = 1; return; }  // converged: last x step pending
      if (st[1] + 1 >= niter) { st[0] = 1; return; }
      break;
    }
    case kOpCgParBeta: {  // slot kCpVau holds (v, Au)
      sc[kCpBcoef] = (sc[kCpVau] * -1.0) / sc[kCpUau];
      break;
    }
    default: break;
  }
}

// ---- PCG vector kernels ----
__global__ void k_pcg_p(int64_t n, const KState* ks, const double* __restrict__ z, double* __restrict__ p) {
  if (ks->st[0]) return;
  const double beta = ks->sc[kBeta];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = p[i] * beta + 1.0 * z[i];
}

__global__ void k_pcg_xr(int64_t n, const KState* ks, const double* __restrict__ p, const double* __restrict__ q,
                         double* __restrict__ x, double* __restrict__ r) {
  if (ks->st[0]) return;
  const double alpha = ks->sc[kAlpha];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    r[i] = r[i] + (-alpha) * q[i];
    x[i] = x[i] + alpha * p[i];
  }
}

// max_i |r_i| (MaxProdSVect, svector.cpp:376-385, with the unit norm vector
// every Abstract*Solver passes, solvers.cpp:5554,5762): per-block maxima,
// then one block folds them into sc[slot]. max is exact in any order.
constexpr int kMaxBlocks = 1024;
__global__ __launch_bounds__(256) void k_maxabs(int64_t n, const int* skip, const double* __restrict__ r,
                                                double* part) {
  if (*skip) return;
  __shared__ double sm[4];
  double m = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) m = fmax(m, fabs(r[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

__global__ __launch_bounds__(256) void k_maxabs_fold(int nb, const int* skip, const double* __restrict__ part,
                                                     KState* ks, int slot) {
  if (*skip) return;
  __shared__ double sm[4];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, part[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) ks->sc[slot] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// ---- BiCGStab vector kernels ----
__global__ void k_bicg_p(int64_t n, const KState* ks, const double* __restrict__ r, const double* __restrict__ v,
                         double* __restrict__ p) {
  if (ks->st[0]) return;
  const double beta = (ks->sc[kRho] / ks->sc[kRhoOld]) * (ks->sc[kBAlpha] / ks->sc[kOmega]);
  const double w = ks->sc[kOmega];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = r[i] + beta * (p[i] - w * v[i]);
}

__global__ __launch_bounds__(256) void k_bicg_s(int64_t n, const KState* ks, const double* __restrict__ r,
                                                const double* __restrict__ v, double* __restrict__ s,
                                                double* partials) {
  if (ks->st[0]) return;
  const double a = ks->sc[kBAlpha];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double si = r[i] - a * v[i];
    s[i] = si;
    acc[0] += si * si;
  }
  block_partials<1>(acc, partials);
}

__global__ void k_bicg_xs(int64_t n, const KState* ks, int stamp, const double* __restrict__ phat,
                          double* __restrict__ x) {
  if (ks->st[3] != stamp) return;  // only in the iteration that converged on s
  const double a = ks->sc[kBAlpha];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] += a * phat[i];
}

__global__ __launch_bounds__(256) void k_bicg_xr(int64_t n, const KState* ks, const double* __restrict__ phat,
                                                 const double* __restrict__ shat, const double* __restrict__ s,
                                                 const double* __restrict__ t, const double* __restrict__ rhat,
                                                 double* __restrict__ x, double* __restrict__ r,
                                                 double* partials) {
  if (ks->st[0]) return;
  const double a = ks->sc[kBAlpha], w = ks->sc[kOmega];
  double acc[2] = {0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x[i] = x[i] + a * phat[i] + w * shat[i];
    const double ri = s[i] - w * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += rhat[i] * ri;
  }
  block_partials<2>(acc, partials);  // r.r and rhat.r, added by the next scalar step
}

// ---- parallel-form CG vector kernels (cg.cpp:355-771; r = A x - b) ----
__global__ void k_cp_sub(int64_t n, double* __restrict__ r, const double* __restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) r[i] = r[i] - b[i];
}

__global__ __launch_bounds__(256) void k_cp_r(int64_t n, const KState* ks, const double* __restrict__ au,
                                              double* __restrict__ r, double* partials) {
  if (ks->st[0]) return;
  const double a = ks->sc[kCpAcoef];
  double acc[1] = {0.0};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ri = r[i] + a * au[i];
    r[i] = ri;
    acc[0] += ri * ri;
  }
  block_partials<1>(acc, partials);
}

__global__ void k_cp_su(int64_t n, const KState* ks, const double* __restrict__ v, double* __restrict__ x,
                        double* __restrict__ u) {
  if (ks->st[0]) return;
  const double a = ks->sc[kCpAcoef], bc = ks->sc[kCpBcoef];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double ui = u[i];
    x[i] = x[i] + a * ui;
    u[i] = v[i] + bc * ui;
  }
}

// The step x += acoef u that the reference skips when it leaves at the test.
__global__ void k_cp_last(int64_t n, const KState* ks, const double* __restrict__ u, double* __restrict__ x) {
  if (!ks->st[6]) return;
  const double a = ks->sc[kCpAcoef];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] = x[i] + a * u[i];
}

__global__ void k_move2(KState* ks, int from, int to) { ks->sc[to] = ks->sc[from]; }

__global__ void k_cp_init(KState* ks, double eps, int sttype) {
  double* sc = ks->sc;
  const double r0 = sc[kCpRr];
  sc[kR0norm2] = r0;
  if (sttype == 1) sc[kCpThr2] = eps * eps * r0;
  else if (sttype == 3) sc[kCpThr2] = eps * eps * sc[kBnorm2];
  else sc[kCpThr2] = eps * eps;
}

int ew_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 1024));
}

// Launches the scalar step `op`, consuming the pending partials of ctx.rw
// into sc[slot0..].
void scalar_step(Context& ctx, int op, double eps2, int ittype, int niter, int slot0, int stamp = 0) {
  RedWork& rw = ctx.rw;
  const int K = rw.pending_k, nb = rw.pending_blocks;
  rw.pending_k = 0;
  hipLaunchKernelGGL(k_scalar, dim3(1), dim3(256), 0, ctx.stream, op, ctx.dstate(), eps2, ittype, niter,
                     rw.partials.as<double>(), nb, K, slot0, stamp);
}

// Marks the partials a solver vector kernel left in ctx.rw as pending.
void set_pending(Context& ctx, int blocks, int k) {
  ctx.rw.pending_blocks = blocks;
  ctx.rw.pending_k = k;
}

}  // namespace

// ---------------------------------------------------------------------------
// Device images

// Full storage of an upper-stored symmetric matrix: the upper blocks are
// copied and expanded on the device; on the host only if a block lies below
// the diagonal.
static void upload_symmetric(const BlockCsr& upper, DevBsr& d) {
  bool uniform = upper.nsup > 0;
  for (int i = 0; i < upper.nsup && uniform; ++i) uniform = upper.bla(i) == upper.bla(0);
  std::vector<int64_t> full_blocks;
  if (!(uniform && fast_block_size(upper.bla(0))) || !dev_bsr_upload_symm_parts({&upper}, d, full_blocks))
    dev_bsr_upload(symm_mtr(upper), kF64, d);
}

void Matrix::to_device(bool with_transpose) {
  if (!has_d) {
    if (symmetric) {
      upload_symmetric(h, d);
      if (with_transpose) { upload_symmetric(h, dt); has_dt = true; }
    } else if (!(cplx && dev_bsr_upload_complex(h, d))) {
      dev_bsr_upload(h, kF64, d);
    }
    has_d = true;
  }
  if (with_transpose && !has_dt) {
    if (symmetric) upload_symmetric(h, dt);
    else dev_bsr_upload(transp_mtr(h), kF64, dt);
    has_dt = true;
  }
}

const Factor::Image& Factor::to_device(int prec, bool with_transpose) {
  if (prec < kF64 || prec > kBF16) throw Error(kErrArg, "unknown factor precision");
  Image& im = img[prec];
  if (!im.has) {
    dev_tri_forward(lower(), prec, im.fwd, cplx);
    dev_tri_backward(u, prec, im.bwd, cplx);
    im.has = true;
  }
  if (with_transpose && !im.has_t) {
    if (kind == 0) {
      dev_tri_forward(u, prec, im.fwdt);
      dev_tri_backward(u, prec, im.bwdt);
    } else {
      dev_tri_forward(u, prec, im.fwdt, cplx);
      dev_tri_backward(l, prec, im.bwdt, cplx);
    }
    im.has_t = true;
  }
  return im;
}

Context::Context() {
  {
    SetupGuard g;
    BCM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  }
  own_stream = true;
  kstate.alloc(sizeof(KState));
  SetupGuard g;
  BCM_HIP(hipMemsetAsync(kstate.p, 0, sizeof(KState), stream));
  BCM_HIP(hipStreamSynchronize(stream));
  BCM_HIP(hipHostMalloc((void**)&host_state, sizeof(KState), hipHostMallocDefault));
  std::memset(host_state, 0, sizeof(KState));
  rw.ensure();
}

Context::~Context() {
  SetupGuard g;
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto& p : pending) { (void)hipEventDestroy(p.e0); (void)hipEventDestroy(p.e1); }
  for (auto e : ev_pool) (void)hipEventDestroy(e);
  if (host_state) (void)hipHostFree(host_state);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

double* Context::vec(int i, size_t nbytes) {
  if ((int)vecs.size() <= i) vecs.resize(i + 1);
  if (vecs[i].bytes < nbytes) vecs[i].alloc(nbytes);
  return vecs[i].as<double>();
}

void Context::check_sweep_error() {
  if (sweep_mode == 1 && sync.error(stream))
    throw Error(kErrInternal, "sync-free sweep: a dependency wait reached its bound");
}

void Context::pull_state() {
  BCM_HIP(hipMemcpyAsync(host_state, kstate.p, sizeof(KState), hipMemcpyDeviceToHost, stream));
  BCM_HIP(hipStreamSynchronize(stream));
}

hipEvent_t Context::prof_begin() {
  hipEvent_t e;
  if (ev_pool.empty()) BCM_HIP(hipEventCreate(&e));
  else { e = ev_pool.back(); ev_pool.pop_back(); }
  BCM_HIP(hipEventRecord(e, stream));
  return e;
}

void Context::prof_end(int kind, hipEvent_t e0, double bytes) {
  hipEvent_t e1;
  if (ev_pool.empty()) BCM_HIP(hipEventCreate(&e1));
  else { e1 = ev_pool.back(); ev_pool.pop_back(); }
  BCM_HIP(hipEventRecord(e1, stream));
  pending.push_back(Pending{kind, e0, e1, bytes});
  if (pending.size() > 4096) prof_flush();
}

namespace {
// acc += sum_c active[c] * part_bytes[c] (one block, fixed order).
__global__ void k_gate_bytes(int np, const int* __restrict__ active, const double* __restrict__ part_bytes,
                             double* acc) {
  __shared__ double sm[256];
  double t = 0.0;
  for (int c = threadIdx.x; c < np; c += blockDim.x) t += active[c] ? part_bytes[c] : 0.0;
  sm[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) s += sm[i];
    *acc += s;
  }
}
}  // namespace

ProfScope::~ProfScope() {
  if (!(c.prof && e0)) return;
  c.prof_end(kind, e0, bytes);
  if (gate_active) {  // after the closing event: not part of the timed span
    hipLaunchKernelGGL(k_gate_bytes, dim3(1), dim3(256), 0, c.stream, nparts, gate_active, part_bytes,
                       c.prof_dev_bytes() + kind);
    (void)hipGetLastError();
  }
}

double* Context::prof_dev_bytes() {
  if (!prof_dev.p) {
    prof_dev.alloc(kProfKinds * sizeof(double));
    BCM_HIP(hipMemsetAsync(prof_dev.p, 0, kProfKinds * sizeof(double), stream));
  }
  return prof_dev.as<double>();
}

void Context::prof_collect() {
  if (!prof_dev.p) return;
  double h[kProfKinds];
  BCM_HIP(hipMemcpyAsync(h, prof_dev.p, sizeof(h), hipMemcpyDeviceToHost, stream));
  BCM_HIP(hipStreamSynchronize(stream));
  for (int k = 0; k < kProfKinds; ++k) prof_acc[k].bytes += h[k];
  BCM_HIP(hipMemsetAsync(prof_dev.p, 0, kProfKinds * sizeof(double), stream));
}

void Context::prof_flush() {
  if (pending.empty()) return;
  BCM_HIP(hipEventSynchronize(pending.back().e1));
  for (auto& p : pending) {
    float ms = 0.f;
    BCM_HIP(hipEventElapsedTime(&ms, p.e0, p.e1));
    prof_acc[p.kind].ms += ms;
    prof_acc[p.kind].launches += 1;
    prof_acc[p.kind].bytes += p.bytes;
    ev_pool.push_back(p.e0);
    ev_pool.push_back(p.e1);
  }
  pending.clear();
}

void precond_sweeps(const DevTri& t1, const DevTri& t2, Context& ctx, const double* r, double* z, int prec, int nrhs,
                    const int* skip, RowGate gate) {
  const int64_t n = t1.n;
  if (prec == kF64) {
    double* tmp = ctx.vec(12, (size_t)n * nrhs * 8);
    {
      ProfScope ps(ctx, kProfSweep, sweep_bytes(t1, nrhs));
      launch_sweep(t1, r, tmp, kF64, nrhs, ctx.stream, skip, ctx.sweep_sync(), gate);
    }
    ProfScope ps(ctx, kProfSweep, sweep_bytes(t2, nrhs));
    launch_sweep(t2, tmp, z, kF64, nrhs, ctx.stream, skip, ctx.sweep_sync(), gate);
  } else {
    float* r32 = reinterpret_cast<float*>(ctx.vec(13, (size_t)n * nrhs * 4));
    float* t32 = reinterpret_cast<float*>(ctx.vec(14, (size_t)n * nrhs * 4));
    launch_convert(n * nrhs, r, kF64, r32, kF32, ctx.stream, skip);
    {
      ProfScope ps(ctx, kProfSweep, sweep_bytes(t1, nrhs));
      launch_sweep(t1, r32, t32, kF32, nrhs, ctx.stream, skip, ctx.sweep_sync(), gate);
    }
    {
      ProfScope ps(ctx, kProfSweep, sweep_bytes(t2, nrhs));
      launch_sweep(t2, t32, r32, kF32, nrhs, ctx.stream, skip, ctx.sweep_sync(), gate);
    }
    launch_convert(n * nrhs, r32, kF32, z, kF64, ctx.stream, skip);
  }
}

void precond_apply(Factor& F, Context& ctx, const double* r, double* z, int prec, bool transpose, int nrhs,
                   const int* skip) {
  const Factor::Image& im = F.to_device(prec, transpose);
  precond_sweeps(transpose ? im.fwdt : im.fwd, transpose ? im.bwdt : im.bwd, ctx, r, z, prec, nrhs, skip);
}

// ---------------------------------------------------------------------------
// Drivers

namespace {

struct TrueResidual {
  double relres = 0.0, bnorm = 0.0;
};

// relres = ||b - A x|| / ||b|| in fp64 on the device; leaves b - A x in r.
TrueResidual true_residual(Matrix& A, Context& ctx, const double* b, const double* x, double* r) {
  const int n = A.d.n;
  KState* ks = ctx.dstate();
  launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, ctx.stream);
  const double* xs[2] = {r, b};
  const double* ys[2] = {r, b};
  launch_multidot(n, 2, xs, ys, &ks->sc[kTrue0], ctx.rw, ctx.stream);
  ctx.pull_state();
  TrueResidual t;
  t.bnorm = std::sqrt(ctx.host_state->sc[kTrue1]);
  t.relres = t.bnorm > 0 ? std::sqrt(ctx.host_state->sc[kTrue0]) / t.bnorm : std::sqrt(ctx.host_state->sc[kTrue0]);
  return t;
}

void pcg(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x, SolveStats& st) {
  const int n = A.d.n;
  const size_t vb = (size_t)n * 8;
  double* r = ctx.vec(0, vb);
  double* z = ctx.vec(1, vb);
  double* p = ctx.vec(2, vb);
  double* q = ctx.vec(3, vb);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[0];
  // ittype 2 compares max|r_i| with |eps| (cg.cpp:241); the others compare squared 2-norms
  const bool maxnorm = opt.ittype == 2;
  const double eps2 = maxnorm ? std::fabs(opt.eps) : opt.eps * opt.eps;
  const int gm = std::min(kMaxBlocks, ew_grid(n));
  double* mpart = maxnorm ? ctx.vec(15, (size_t)kMaxBlocks * 8) : nullptr;
  hipStream_t s = ctx.stream;
  int total = 0;
  for (int attempt = 0; attempt <= opt.max_restarts; ++attempt) {
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, ks, attempt > 0 ? 1 : 0);
    launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);  // r = b - A x
    launch_fill(n, 0.0, p, s);
    {
      const double* xs[2] = {b, r};
      const double* ys[2] = {b, r};
      launch_multidot(n, 2, xs, ys, &ks->sc[kBnorm2], ctx.rw, s);  // sc[20] = b.b, sc[21] = r0.r0
    }
    int it = 0;
    const bool exact = opt.check_every == 1;
    while (true) {
      for (int c = 0; c < opt.check_every; ++c) {
        precond_apply(F, ctx, r, z, opt.prec, false, 1, done);
        const double* xs[2] = {r, r};
        const double* ys[2] = {r, z};
        launch_multidot(n, 2, xs, ys, kDeferred, ctx.rw, s, done);
        if (maxnorm) {
          hipLaunchKernelGGL(k_maxabs, dim3(gm), dim3(256), 0, s, (int64_t)n, done, r, mpart);
          hipLaunchKernelGGL(k_maxabs_fold, dim3(1), dim3(256), 0, s, gm, done, mpart, ks, (int)kErrloc);
        }
        scalar_step(ctx, kOpPcgCheck, eps2, opt.ittype, opt.niter - total, kEnrsq);  // (r,r) -> 8, (r,z) -> 9
        if (exact) {  // no launch is queued past the convergence test
          ctx.pull_state();
          if (ctx.host_state->st[0]) break;
        }
        hipLaunchKernelGGL(k_pcg_p, dim3(ew_grid(n)), dim3(256), 0, s, (int64_t)n, ks, z, p);
        {
          ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
          launch_spmv(A.d, p, q, nullptr, 1, kDeferred, p, ctx.rw, s, done);
        }
        scalar_step(ctx, kOpPcgAlpha, eps2, opt.ittype, opt.niter - total, kPq);
        hipLaunchKernelGGL(k_pcg_xr, dim3(ew_grid(n)), dim3(256), 0, s, (int64_t)n, ks, p, q, x, r);
        ++it;
      }
      ctx.pull_state();
      if (ctx.host_state->st[0]) break;
    }
    const KState& hs = *ctx.host_state;
    const int k = hs.st[1];
    total += std::max(0, k - 1);
    st.breakdown = hs.st[2];
    st.relres_rec = hs.sc[kBnorm2] > 0 ? std::sqrt(hs.sc[kEnrsq] / hs.sc[kBnorm2]) : std::sqrt(hs.sc[kEnrsq]);
    st.errmin = maxnorm ? hs.sc[kErrloc] : 0.0;
    TrueResidual tr = true_residual(A, ctx, b, x, r);
    st.relres_true = tr.relres;
    const bool ok = opt.ittype == 3 ? tr.relres <= opt.eps : hs.st[4] != 0;
    st.converged = ok ? 1 : 0;
    if (ok || st.breakdown || total >= opt.niter) break;
    st.restarts++;
  }
  st.iterations = total;
}

// Parallel-form CG (CSVector::Cg with operator plug-ins, cg.cpp:355-771) on one
// device: the (Au,u), (r,u) pair and (v,Au) are the reference's reductions
// (its 2-element and 1-element allreduces). Stopping: ittype 0 -> ||r|| <=
// eps ||r0|| (sttype 1), 1 -> ||r|| <= eps (sttype 0), 3 -> ||r|| <= eps ||b||.
// At a converged exit the reference has not yet added the last acoef u to x
// (cg.cpp:607-614 against :660-661); this driver adds it, so the returned x
// is the iterate whose residual met the test.
void cg_par(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
            SolveStats& st) {
  const int64_t n = A.d.n;
  const size_t vb = (size_t)n * 8;
  double* r = ctx.vec(0, vb);
  double* u = ctx.vec(1, vb);
  double* au = ctx.vec(2, vb);
  double* v = ctx.vec(3, vb);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[0];
  hipStream_t s = ctx.stream;
  const int ge = ew_grid(n);
  const int sttype = opt.ittype == 0 ? 1 : opt.ittype == 3 ? 3 : 0;
  hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, ks, 0);
  launch_spmv(A.d, x, r, nullptr, 1, nullptr, nullptr, ctx.rw, s);
  hipLaunchKernelGGL(k_cp_sub, dim3(ge), dim3(256), 0, s, n, r, b);
  {
    const double* xs[2] = {r, b};
    launch_multidot((int)n, 2, xs, xs, &ks->sc[kCpRr], ctx.rw, s);  // kCpRr: r.r, kCpVau(16): b.b
    hipLaunchKernelGGL(k_move2, dim3(1), dim3(1), 0, s, ks, kCpVau, kBnorm2);
    hipLaunchKernelGGL(k_cp_init, dim3(1), dim3(1), 0, s, ks, opt.eps, sttype);
  }
  precond_apply(F, ctx, r, u, opt.prec, false, 1);
  const bool exact = opt.check_every == 1;
  while (true) {
    for (int c = 0; c < std::max(1, opt.check_every); ++c) {
      {
        ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
        launch_spmv(A.d, u, au, nullptr, 1, nullptr, nullptr, ctx.rw, s, done);
      }
      {
        const double* xs[2] = {au, r};
        const double* ys[2] = {u, u};
        launch_multidot((int)n, 2, xs, ys, kDeferred, ctx.rw, s, done);
      }
      scalar_step(ctx, kOpCgParAlpha, 0.0, opt.ittype, opt.niter, kCpUau);
      hipLaunchKernelGGL(k_cp_r, dim3(ge), dim3(256), 0, s, n, ks, au, r, ctx.rw.partials.as<double>());
      set_pending(ctx, ge, 1);
      scalar_step(ctx, kOpCgParCheck, 0.0, opt.ittype, opt.niter, kCpRr);
      precond_apply(F, ctx, r, v, opt.prec, false, 1, done);
      {
        const double* xs[1] = {v};
        const double* ys[1] = {au};
        launch_multidot((int)n, 1, xs, ys, kDeferred, ctx.rw, s, done);
      }
      scalar_step(ctx, kOpCgParBeta, 0.0, opt.ittype, opt.niter, kCpVau);
      hipLaunchKernelGGL(k_cp_su, dim3(ge), dim3(256), 0, s, n, ks, v, x, u);
      if (exact) {
        ctx.pull_state();
        if (ctx.host_state->st[0]) break;
      }
    }
    ctx.pull_state();
    if (ctx.host_state->st[0]) break;
  }
  hipLaunchKernelGGL(k_cp_last, dim3(ge), dim3(256), 0, s, n, ks, u, x);
  ctx.pull_state();
  const KState hs = *ctx.host_state;
  st.iterations = hs.st[1];
  st.breakdown = hs.st[2];
  st.relres_rec = hs.sc[kBnorm2] > 0 ? std::sqrt(hs.sc[kCpRr] / hs.sc[kBnorm2]) : std::sqrt(hs.sc[kCpRr]);
  TrueResidual tr = true_residual(A, ctx, b, x, r);
  st.relres_true = tr.relres;
  st.converged = hs.st[4] ? 1 : 0;
}

void bicgstab(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
              SolveStats& st) {
  const int n = A.d.n;
  const size_t vb = (size_t)n * 8;
  double* r = ctx.vec(0, vb);
  double* rhat = ctx.vec(1, vb);
  double* p = ctx.vec(2, vb);
  double* v = ctx.vec(3, vb);
  double* phat = ctx.vec(4, vb);
  double* sv = ctx.vec(5, vb);
  double* shat = ctx.vec(6, vb);
  double* t = ctx.vec(7, vb);
  KState* ks = ctx.dstate();
  const int* done = &ks->st[0];
  const double eps2 = opt.eps * opt.eps;
  hipStream_t s = ctx.stream;
  const int ge = ew_grid(n);
  int total = 0;
  int stamp = 0;  // identifies each queued iteration (k_bicg_xs acts only for the one that converged on s)
  for (int attempt = 0; attempt <= opt.max_restarts; ++attempt) {
    hipLaunchKernelGGL(k_state_reset, dim3(1), dim3(64), 0, s, ks, 0);
    launch_spmv(A.d, x, r, b, 1, nullptr, nullptr, ctx.rw, s);
    launch_copy(n, r, rhat, s);
    launch_fill(n, 0.0, p, s);
    launch_fill(n, 0.0, v, s);
    {
      const double* xs[3] = {b, r, rhat};
      const double* ys[3] = {b, r, r};
      launch_multidot(n, 3, xs, ys, kDeferred, ctx.rw, s);
      scalar_step(ctx, kOpBicgInit, eps2, opt.ittype, opt.niter, kBnorm2);  // 20: b.b 21: r.r 22: rhat.r
    }
    const bool exact = opt.check_every == 1;
    while (true) {
      for (int c = 0; c < opt.check_every; ++c) {
        ++stamp;
        hipLaunchKernelGGL(k_bicg_p, dim3(ge), dim3(256), 0, s, (int64_t)n, ks, r, v, p);
        precond_apply(F, ctx, p, phat, opt.prec, false, 1, done);
        {
          ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
          launch_spmv(A.d, phat, v, nullptr, 1, kDeferred, rhat, ctx.rw, s, done);
        }
        scalar_step(ctx, kOpBicgAlpha, eps2, opt.ittype, opt.niter, kRv);
        hipLaunchKernelGGL(k_bicg_s, dim3(ge), dim3(256), 0, s, (int64_t)n, ks, r, v, sv, ctx.rw.partials.as<double>());
        set_pending(ctx, ge, 1);
        scalar_step(ctx, kOpBicgSCheck, eps2, opt.ittype, opt.niter, kSs, stamp);
        hipLaunchKernelGGL(k_bicg_xs, dim3(ge), dim3(256), 0, s, (int64_t)n, ks, stamp, phat, x);
        if (exact) {  // no launch is queued past a convergence test
          ctx.pull_state();
          if (ctx.host_state->st[0]) break;
        }
        precond_apply(F, ctx, sv, shat, opt.prec, false, 1, done);
        {
          ProfScope ps(ctx, kProfSpmv, spmv_bytes(A.d, 1));
          launch_spmv(A.d, shat, t, nullptr, 1, nullptr, nullptr, ctx.rw, s, done);
        }
        {
          const double* xs[2] = {t, t};
          const double* ys[2] = {sv, t};
          launch_multidot(n, 2, xs, ys, kDeferred, ctx.rw, s, done);
        }
        scalar_step(ctx, kOpBicgOmega, eps2, opt.ittype, opt.niter, kTs);
        hipLaunchKernelGGL(k_bicg_xr, dim3(ge), dim3(256), 0, s, (int64_t)n, ks, phat, shat, sv, t, rhat, x, r,
                           ctx.rw.partials.as<double>());
        set_pending(ctx, ge, 2);
        scalar_step(ctx, kOpBicgNext, eps2, opt.ittype, opt.niter - total, kRr);
        if (exact) {
          ctx.pull_state();
          if (ctx.host_state->st[0]) break;
        }
      }
      ctx.pull_state();
      if (ctx.host_state->st[0]) break;
    }
    const KState& hs = *ctx.host_state;
    total += hs.st[1];
    st.breakdown = hs.st[2];
    st.relres_rec = hs.sc[kBnorm2] > 0 ? std::sqrt((hs.st[3] ? hs.sc[kSs] : hs.sc[kRr]) / hs.sc[kBnorm2]) : 0.0;
    TrueResidual tr = true_residual(A, ctx, b, x, r);
    st.relres_true = tr.relres;
    const bool ok = opt.ittype == 3 ? tr.relres <= opt.eps : hs.st[4] != 0;
    st.converged = ok ? 1 : 0;
    if (ok || total >= opt.niter) break;
    st.restarts++;
  }
  st.iterations = total;
}

}  // namespace

void solve(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
           SolveStats& st) {
  auto t0 = std::chrono::steady_clock::now();
  st = SolveStats();
  A.to_device(opt.kind == kLanczos || blz_variant(opt.kind) >= 0);
  if (A.d.n != F.u.n()) throw Error(kErrArg, "matrix and factor sizes differ");
  auto run = [&] {
    switch (opt.kind) {
      case kPcg: pcg(A, F, ctx, opt, b, x, st); break;
      case kBicgstab: bicgstab(A, F, ctx, opt, b, x, st); break;
      case kCgPar: cg_par(A, F, ctx, opt, b, x, st); break;
      case kFgmres: solve_fgmres(A, F, ctx, opt, b, x, st); break;
      case kLanczos: solve_lanczos(A, F, ctx, opt, b, x, st); break;
      case kBlockGmres: solve_block_gmres(A, F, ctx, opt, 1, b, x, st); break;
      case kBlzCentral:
      case kBlzRight:
      case kBlzLeft: solve_block_lanczos(A, F, ctx, opt, blz_variant(opt.kind), 1, b, x, st); break;
      default: throw Error(kErrArg, "solver kind not available");
    }
    BCM_HIP(hipStreamSynchronize(ctx.stream));
  };
  // Sync-free sweeps are first launched as ordinary kernels: the grid never
  // exceeds what the device holds at once, so on an otherwise idle GPU every
  // wave is resident, without the cooperative launch's cost. If a wait hit
  // its bound (another process held part of the GPU), the solve is repeated
  // from the same x0 with cooperative launches, which guarantee residency.
  const bool saved = ctx.sync.cooperative;
  const bool fast = ctx.sweep_mode == 1 && saved;
  const int64_t n = A.d.n;
  double* x0 = nullptr;
  if (fast) {
    x0 = ctx.vec(87, (size_t)n * 8);
    launch_copy(n, x, x0, ctx.stream);
    ctx.sync.cooperative = false;
  }
  try {
    run();
  } catch (...) {
    ctx.sync.cooperative = saved;
    throw;
  }
  ctx.sync.cooperative = saved;
  if (fast && ctx.sync.error(ctx.stream)) {
    ctx.sync.clear(ctx.stream);
    launch_copy(n, x0, x, ctx.stream);
    st = SolveStats();
    run();
  }
  ctx.check_sweep_error();
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace bcm
