// Initial-guess model (Guess): snapshot store, POD bases by a randomized
// range finder on the device, and the MLP that maps cell parameters to POD
// coefficients. Host code only; the device work is in learn.hip.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <random>

#include "learn.hpp"

namespace bcm {

namespace {

// Upper-triangular R with G = R^T R (G r x r column-major, SPD); returns R^-1.
std::vector<double> chol_inv_upper(int r, const std::vector<double>& G) {
  std::vector<double> R((size_t)r * r, 0.0);
  for (int j = 0; j < r; ++j) {
    for (int i = 0; i <= j; ++i) {
      double s = G[(size_t)j * r + i];
      for (int k = 0; k < i; ++k) s -= R[(size_t)i * r + k] * R[(size_t)j * r + k];
      if (i == j) {
        if (!(s > 0.0)) s = 1e-300;  // rank-deficient snapshots: keep the column, it is rescaled to zero weight
        R[(size_t)j * r + j] = std::sqrt(s);
      } else {
        R[(size_t)j * r + i] = s / R[(size_t)i * r + i];
      }
    }
  }
  // inverse of the upper-triangular R (column-major), column by column
  std::vector<double> Ri((size_t)r * r, 0.0);
  for (int j = 0; j < r; ++j) {
    Ri[(size_t)j * r + j] = 1.0 / R[(size_t)j * r + j];
    for (int i = j - 1; i >= 0; --i) {
      double s = 0.0;
      for (int k = i + 1; k <= j; ++k) s += R[(size_t)k * r + i] * Ri[(size_t)j * r + k];
      Ri[(size_t)j * r + i] = -s / R[(size_t)i * r + i];
    }
  }
  return Ri;
}

template <class T>
void to_dev(DevBuf& d, const std::vector<T>& h, hipStream_t s) {
  const size_t nb = h.size() * sizeof(T);
  if (d.bytes < nb) d.alloc(nb);
  BCM_HIP(hipMemcpyAsync(d.p, h.data(), nb, hipMemcpyHostToDevice, s));
  BCM_HIP(hipStreamSynchronize(s));
}

template <class T>
std::vector<T> to_host(const void* d, size_t count, hipStream_t s) {
  std::vector<T> h(count);
  BCM_HIP(hipMemcpyAsync(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost, s));
  BCM_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace

Guess::Guess(int64_t n_, int nrhs_, int nparam_, int nmodes_, int nh_, int capacity, uint64_t seed_)
    : n(n_), nrhs(nrhs_), nparam(nparam_), nmodes(nmodes_), nh(nh_), cap(capacity), seed(seed_) {
  if (n < 1 || nrhs < 1 || nrhs > 16 || nparam < 1 || nparam > 64 || nmodes < 1 || nmodes > 256 || capacity < 1)
    throw Error(kErrArg, "guess: bad sizes");
  snaps.resize(nrhs);
  for (auto& b : snaps) b.alloc((size_t)n * cap * 8);
}

void Guess::add(hipStream_t s, int cnt, const double* prm, const double* x) {
  if (cnt < 1 || count + cnt > cap) throw Error(kErrArg, "guess: snapshot capacity exceeded");
  for (int q = 0; q < nrhs; ++q)
    BCM_HIP(hipMemcpyAsync(snaps[q].as<double>() + (size_t)count * n, x + (size_t)q * n * cnt,
                           (size_t)n * cnt * 8, hipMemcpyDeviceToDevice, s));
  BCM_HIP(hipStreamSynchronize(s));
  params.insert(params.end(), prm, prm + (size_t)cnt * nparam);
  count += cnt;
  fitted = false;
}

GuessReport Guess::fit(hipStream_t s, int steps, double lr) {
  if (count < 2) throw Error(kErrArg, "guess: at least two snapshots are needed");
  auto t0 = std::chrono::steady_clock::now();
  GuessReport rep;
  rep.snapshots = count;
  const int k = count;
  const int modes = std::min(nmodes, k);
  const int r = std::min(k, modes + 16);  // oversampled range
  rep.modes = modes;
  basis.resize(nrhs);
  std::vector<std::vector<double>> coef(nrhs);  // per rhs: modes x k, column-major
  std::mt19937_64 g(seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  DevBuf om, z, G;
  work1.alloc((size_t)n * r * 8);
  work2.alloc((size_t)n * r * 8);
  scratch.alloc((size_t)64 * k * r * 8);  // largest split-n partials of tsmm: sized once, never resized in flight
  double energy_sum = 0.0;
  for (int q = 0; q < nrhs; ++q) {
    const double* S = snaps[q].as<const double>();
    double* Y = work1.as<double>();
    double* Y2 = work2.as<double>();
    std::vector<double> omega((size_t)k * r);
    for (auto& v : omega) v = nd(g);
    to_dev(om, omega, s);
    if (z.bytes < (size_t)k * r * 8) z.alloc((size_t)k * r * 8);
    if (G.bytes < (size_t)std::max(r * r, r * k) * 8) G.alloc((size_t)std::max(r * r, r * k) * 8);
    mm(s, n, k, r, S, om.as<double>(), Y, n);  // Y = S Omega
    auto cholqr = [&]() {                     // Y <- Y R^-1 with Y^T Y = R^T R
      tsmm(s, n, r, r, Y, Y, G.as<double>(), scratch);
      std::vector<double> gh = to_host<double>(G.p, (size_t)r * r, s);
      std::vector<double> ri = chol_inv_upper(r, gh);
      to_dev(om, ri, s);
      mm(s, n, r, r, Y, om.as<double>(), Y2, n);
      std::swap(Y, Y2);
    };
    for (int it = 0; it < 2; ++it) {  // power iterations
      cholqr();
      tsmm(s, n, k, r, S, Y, z.as<double>(), scratch);  // Z = S^T Y (k x r)
      mm(s, n, k, r, S, z.as<double>(), Y, n);          // Y = S Z
    }
    cholqr();
    cholqr();
    // B = Q^T S (r x k); small eigenproblem B B^T = W diag(sigma^2) W^T on the host
    tsmm(s, n, r, k, Y, S, G.as<double>(), scratch);
    std::vector<double> bh = to_host<double>(G.p, (size_t)r * k, s);
    std::vector<double> mb((size_t)r * r, 0.0);
    for (int j = 0; j < r; ++j)
      for (int i = 0; i <= j; ++i) {
        double acc = 0.0;
        for (int c = 0; c < k; ++c) acc += bh[(size_t)c * r + i] * bh[(size_t)c * r + j];
        mb[(size_t)j * r + i] = mb[(size_t)i * r + j] = acc;
      }
    double total = 0.0;
    for (int i = 0; i < r; ++i) total += mb[(size_t)i * r + i];
    std::vector<double> ev(r);
    sym_eig(r, mb.data(), ev.data());  // ascending
    std::vector<double> w((size_t)r * modes);
    double kept = 0.0;
    for (int m = 0; m < modes; ++m) {
      const int src = r - 1 - m;
      kept += std::max(0.0, ev[src]);
      std::memcpy(&w[(size_t)m * r], &mb[(size_t)src * r], (size_t)r * 8);
    }
    energy_sum += total > 0.0 ? kept / total : 1.0;
    to_dev(om, w, s);
    basis[q].alloc((size_t)n * modes * 8);
    mm(s, n, r, modes, Y, om.as<double>(), basis[q].as<double>(), n);  // Phi = Q W
    // training coefficients: Phi^T S (modes x k)
    tsmm(s, n, modes, k, basis[q].as<const double>(), S, G.as<double>(), scratch);
    coef[q] = to_host<double>(G.p, (size_t)modes * k, s);
  }
  rep.pod_energy = energy_sum / nrhs;
  auto t1 = std::chrono::steady_clock::now();
  rep.seconds_pod = std::chrono::duration<double>(t1 - t0).count();

  // normalisation: inputs to zero mean / unit spread; outputs centred, one
  // scale per right-hand side (so the loss stays proportional to ||x0 - x||^2)
  const int nout = nrhs * modes;
  in_mean.assign(nparam, 0.0);
  in_scale.assign(nparam, 1.0);
  for (int j = 0; j < nparam; ++j) {
    double m = 0.0, v = 0.0;
    for (int i = 0; i < k; ++i) m += params[(size_t)i * nparam + j];
    m /= k;
    for (int i = 0; i < k; ++i) v += (params[(size_t)i * nparam + j] - m) * (params[(size_t)i * nparam + j] - m);
    v = std::sqrt(v / k);
    in_mean[j] = m;
    in_scale[j] = v > 0 ? v : 1.0;
  }
  out_mean.assign(nout, 0.0);
  out_scale.assign(nout, 1.0);
  for (int q = 0; q < nrhs; ++q) {
    double ss = 0.0;
    for (int m = 0; m < modes; ++m) {
      double mu = 0.0;
      for (int i = 0; i < k; ++i) mu += coef[q][(size_t)i * modes + m];
      mu /= k;
      out_mean[q * modes + m] = mu;
      for (int i = 0; i < k; ++i) ss += (coef[q][(size_t)i * modes + m] - mu) * (coef[q][(size_t)i * modes + m] - mu);
    }
    const double sc = std::sqrt(ss / ((double)k * modes));
    for (int m = 0; m < modes; ++m) out_scale[q * modes + m] = sc > 0 ? sc : 1.0;
  }
  std::vector<float> xin_h((size_t)k * nparam), t_h((size_t)k * nout);
  for (int i = 0; i < k; ++i) {
    for (int j = 0; j < nparam; ++j)
      xin_h[(size_t)i * nparam + j] = (float)((params[(size_t)i * nparam + j] - in_mean[j]) / in_scale[j]);
    for (int q = 0; q < nrhs; ++q)
      for (int m = 0; m < modes; ++m) {
        const int o = q * modes + m;
        t_h[(size_t)i * nout + o] = (float)((coef[q][(size_t)i * modes + m] - out_mean[o]) / out_scale[o]);
      }
  }
  to_dev(xin, xin_h, s);
  to_dev(tout, t_h, s);
  if (loss.bytes < 16) loss.alloc(16);
  mlp.reset(new Mlp(nparam, nh, nout, seed ^ 0x5DEECE66Dull));
  nmodes = modes;
  const float* X = xin.as<const float>();
  const float* T = tout.as<const float>();
  for (int it = 0; it < steps; ++it) {
    const bool rec = it == 0 || it == steps - 1;
    mlp->backward(s, k, X, T, rec ? loss.as<double>() : nullptr);
    if (rec) {
      std::vector<double> l = to_host<double>(loss.p, 1, s);
      (it == 0 ? rep.loss_first : rep.loss_final) = l[0];
    }
    // cosine decay from lr to lr / 20
    const double f = 0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.141592653589793 * it / std::max(1, steps)));
    mlp->adam(s, (float)(lr * f));
  }
  if (steps == 0) rep.loss_first = rep.loss_final = 0.0;
  BCM_HIP(hipStreamSynchronize(s));
  rep.steps = steps;
  rep.seconds_train = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
  fitted = true;
  return rep;
}

void Guess::predict(hipStream_t s, int cnt, const double* prm, double* x0) {
  if (!fitted) throw Error(kErrArg, "guess: fit() has not run");
  if (cnt < 1) throw Error(kErrArg, "guess: no cells");
  const int nout = nrhs * nmodes;
  std::vector<float> xin_h((size_t)cnt * nparam);
  for (int i = 0; i < cnt; ++i)
    for (int j = 0; j < nparam; ++j)
      xin_h[(size_t)i * nparam + j] = (float)((prm[(size_t)i * nparam + j] - in_mean[j]) / in_scale[j]);
  DevBuf xi, yo, cq;
  to_dev(xi, xin_h, s);
  yo.alloc((size_t)cnt * nout * 4);
  mlp->forward(s, cnt, xi.as<const float>(), yo.as<float>());
  std::vector<float> y = to_host<float>(yo.p, (size_t)cnt * nout, s);
  std::vector<double> c((size_t)nmodes * cnt);
  for (int q = 0; q < nrhs; ++q) {
    for (int i = 0; i < cnt; ++i)
      for (int m = 0; m < nmodes; ++m) {
        const int o = q * nmodes + m;
        c[(size_t)i * nmodes + m] = (double)y[(size_t)i * nout + o] * out_scale[o] + out_mean[o];
      }
    to_dev(cq, c, s);
    mm(s, n, nmodes, cnt, basis[q].as<const double>(), cq.as<const double>(), x0 + (size_t)q * n * cnt, n);
    BCM_HIP(hipStreamSynchronize(s));
  }
}

}  // namespace bcm
