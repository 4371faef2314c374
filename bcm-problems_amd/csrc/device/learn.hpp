// Learned initial guess: MLP (BF16 MFMA forward/backward, Adam) and the POD
// bases whose coefficients it predicts. No counterpart in the reference:
// the reference solves every cell from the caller's initial guess.
#pragma once

#include <memory>
#include <vector>

#include "device.hpp"

namespace bcm {

// nin -> nh (tanh) -> nh (tanh) -> nout. Row-major activations (batch x
// features); parameters in one FP32 array [W1 b1 W2 b2 W3 b3] with W_l
// (fan_in x fan_out, row-major), plus BF16 copies of the weights.
struct Mlp {
  int nin, nh, nout;
  int64_t nparam = 0, nweight = 0;
  int64_t w_off[3], b_off[3], wb_off[3];
  DevBuf params, grads, m1, m2, wbf, is_w, wb_index;
  DevBuf h1, h2, y, dz1, dz2, dy, partials, colpart, kpart;
  int cap_batch = 0;
  int step = 0;  // Adam steps taken
  Mlp(int nin, int nh, int nout, uint64_t seed);
  void set_params(const float* host);  // also resets the Adam moments
  void ensure_batch(int M);
  // out (M x nout) = MLP(x); out == nullptr keeps the result in y.
  void forward(hipStream_t s, int M, const float* x, float* out);
  // Gradients of mean((MLP(x) - t)^2) into grads; loss_dev (may be null) receives the loss.
  void backward(hipStream_t s, int M, const float* x, const float* t, double* loss_dev);
  void adam(hipStream_t s, float lr);
};

// C (p x q) = X^T Y for tall X (n x p), Y (n x q), column-major with leading
// dimension n, FP64; deterministic split-n partial sums.
void tsmm(hipStream_t s, int64_t n, int p, int q, const double* X, const double* Y, double* C, DevBuf& scratch);
// Y (n x q, leading dimension ldy) = X (n x p) C (p x q), FP64.
void mm(hipStream_t s, int64_t n, int p, int q, const double* X, const double* Cm, double* Y, int64_t ldy);

struct GuessReport {
  int snapshots = 0, modes = 0, steps = 0;
  double loss_first = 0.0, loss_final = 0.0;
  double pod_energy = 0.0;  // fraction of the snapshot energy the bases keep (mean over right-hand sides)
  double seconds_pod = 0.0, seconds_train = 0.0;
};

// Initial-guess model for a family of cells with one layout (n unknowns,
// nrhs right-hand sides, nparam physics parameters each).
struct Guess {
  int64_t n;
  int nrhs, nparam, nmodes, nh;
  int cap = 0, count = 0;          // snapshot capacity and count
  std::vector<DevBuf> snaps;       // per right-hand side: n x cap, column-major
  std::vector<double> params;      // count x nparam
  std::vector<DevBuf> basis;       // per right-hand side: n x nmodes
  std::unique_ptr<Mlp> mlp;
  std::vector<double> in_mean, in_scale, out_mean, out_scale;
  bool fitted = false;
  DevBuf scratch, work1, work2, work3, xin, tout, loss;
  uint64_t seed;
  Guess(int64_t n, int nrhs, int nparam, int nmodes, int nh, int capacity, uint64_t seed);
  // Appends `cnt` solved cells: x (device) in batch layout, column q holding
  // the cells' q-th solutions one after another (n each).
  void add(hipStream_t s, int cnt, const double* prm, const double* x);
  GuessReport fit(hipStream_t s, int steps, double lr);
  // x0 (device, batch layout of cnt cells) from the cells' parameters.
  void predict(hipStream_t s, int cnt, const double* prm, double* x0);
};

}  // namespace bcm
