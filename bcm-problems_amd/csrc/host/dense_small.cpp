// Small dense eigen/singular-value decompositions used by the IC2/ILU2 host
// factorizations for the scaling of every diagonal block and for the block
// pivots (the reference calls LAPACK 3.0 DSYEV/DGESVD through lapackd.cpp:865-905).
//
// sym_eig: Householder reduction to tridiagonal form followed by the implicit
//   QL iteration with Wilkinson-type shifts (Wilkinson & Reinsch, Handbook for
//   Automatic Computation II, procedures tred2/tql2; EISPACK).
// svd_square: one-sided (Hestenes) Jacobi SVD with the cyclic pivot order
//   (Demmel & Veselic 1992), which reaches full relative accuracy.
//
// Both return the LAPACK conventions the reference relies on: eigenvalues
// ascending with eigenvectors in the columns; singular values descending with
// U and V^T. Eigen/singular vectors are unique up to sign for simple
// eigen/singular values, so the IC2/ILU2 drop decisions (max-abs tests) and the
// preconditioner they define do not depend on the routine used.
#include "bcm_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace bcm {

namespace {

// Householder tridiagonalisation of the symmetric matrix z (row-major view,
// n x n); on exit z holds the accumulated orthogonal transform, d the
// diagonal and e the sub-diagonal (e[0] = 0).
void tridiagonalize(int n, double* z, double* d, double* e) {
  auto Z = [&](int i, int j) -> double& { return z[i * n + j]; };
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
      if (scale == 0.0) {
        e[i] = Z(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          Z(i, k) /= scale;
          h += Z(i, k) * Z(i, k);
        }
        double f = Z(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        Z(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          Z(j, i) = Z(i, j) / h;
          double gg = 0.0;
          for (int k = 0; k <= j; ++k) gg += Z(j, k) * Z(i, k);
          for (int k = j + 1; k <= l; ++k) gg += Z(k, j) * Z(i, k);
          e[j] = gg / h;
          f += e[j] * Z(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          double ff = Z(i, j);
          double gg = e[j] - hh * ff;
          e[j] = gg;
          for (int k = 0; k <= j; ++k) Z(j, k) -= ff * e[k] + gg * Z(i, k);
        }
      }
    } else {
      e[i] = Z(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += Z(i, k) * Z(k, j);
        for (int k = 0; k <= l; ++k) Z(k, j) -= g * Z(k, i);
      }
    }
    d[i] = Z(i, i);
    Z(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0.0;
  }
}

// Implicit QL iteration on the tridiagonal (d, e); rotations accumulated into
// the rows of zt (zt = Z^T, so every rotation updates two contiguous rows).
void tridiagonal_ql(int n, double* d, double* e, double* zt) {
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) throw Error(kErrInternal, " Error in the symmetric eigensolver (no convergence)");
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        bool deflated = false;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            deflated = true;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          double* zi = zt + (size_t)i * n;
          double* zi1 = zt + (size_t)(i + 1) * n;
          for (int k = 0; k < n; ++k) {
            const double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        }
        if (deflated) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
}

}  // namespace

void sym_eig(int n, double* a, double* w) {
  if (n == 1) {
    w[0] = a[0];
    a[0] = 1.0;
    return;
  }
  // a is column-major symmetric: the row-major view is its transpose, i.e. the same matrix.
  std::vector<double> z(a, a + n * n), zt((size_t)n * n), e(n);
  tridiagonalize(n, z.data(), w, e.data());
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) zt[(size_t)i * n + k] = z[(size_t)k * n + i];
  tridiagonal_ql(n, w, e.data(), zt.data());
  // Sort ascending; eigenvector i is row i of zt.
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return w[x] < w[y]; });
  std::vector<double> ws(w, w + n);
  for (int j = 0; j < n; ++j) {
    w[j] = ws[idx[j]];
    std::memcpy(a + (size_t)j * n, zt.data() + (size_t)idx[j] * n, sizeof(double) * n);
  }
}

void svd_square(int n, const double* a, double* s, double* u, double* vt) {
  std::vector<double> g(a, a + n * n);  // columns are transformed in place: G = A V
  std::vector<double> v(n * n, 0.0);
  for (int i = 0; i < n; ++i) v[i * n + i] = 1.0;
  const double tol = 1.0e-15;
  for (int sweep = 0; sweep < 60; ++sweep) {
    bool rotated = false;
    for (int p = 0; p < n - 1; ++p) {
      for (int q = p + 1; q < n; ++q) {
        double* gp = &g[p * n];
        double* gq = &g[q * n];
        double alpha = 0.0, beta = 0.0, gamma = 0.0;
        for (int k = 0; k < n; ++k) {
          alpha += gp[k] * gp[k];
          beta += gq[k] * gq[k];
          gamma += gp[k] * gq[k];
        }
        if (gamma == 0.0 || std::fabs(gamma) <= tol * std::sqrt(alpha * beta)) continue;
        rotated = true;
        double zeta = (beta - alpha) / (2.0 * gamma);
        double t = std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t);
        double sn = c * t;
        for (int k = 0; k < n; ++k) {
          double x = gp[k], y = gq[k];
          gp[k] = c * x - sn * y;
          gq[k] = sn * x + c * y;
        }
        double* vp = &v[p * n];
        double* vq = &v[q * n];
        for (int k = 0; k < n; ++k) {
          double x = vp[k], y = vq[k];
          vp[k] = c * x - sn * y;
          vq[k] = sn * x + c * y;
        }
      }
    }
    if (!rotated) break;
  }
  std::vector<double> sv(n);
  for (int j = 0; j < n; ++j) {
    double t = 0.0;
    for (int k = 0; k < n; ++k) t += g[j * n + k] * g[j * n + k];
    sv[j] = std::sqrt(t);
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return sv[x] > sv[y]; });
  const double smax = sv[idx[0]];
  int rank = 0;
  for (int j = 0; j < n; ++j) {
    int c = idx[j];
    s[j] = sv[c];
    for (int k = 0; k < n; ++k) vt[k * n + j] = v[c * n + k];  // V^T(j, k) = V(k, j)
    if (sv[c] > 1e-300 && sv[c] > smax * 1e-300) {
      for (int k = 0; k < n; ++k) u[j * n + k] = g[c * n + k] / sv[c];
      ++rank;
    } else {
      for (int k = 0; k < n; ++k) u[j * n + k] = 0.0;
    }
  }
  // Complete U to an orthonormal basis where singular values vanished.
  for (int j = rank; j < n; ++j) {
    for (int trial = 0; trial < n; ++trial) {
      double* col = &u[j * n];
      for (int k = 0; k < n; ++k) col[k] = (k == (j + trial) % n) ? 1.0 : 0.0;
      for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < j; ++i) {
          double d = 0.0;
          for (int k = 0; k < n; ++k) d += u[i * n + k] * col[k];
          for (int k = 0; k < n; ++k) col[k] -= d * u[i * n + k];
        }
      double nrm = 0.0;
      for (int k = 0; k < n; ++k) nrm += col[k] * col[k];
      nrm = std::sqrt(nrm);
      if (nrm > 1e-8) {
        for (int k = 0; k < n; ++k) col[k] /= nrm;
        break;
      }
    }
  }
}

}  // namespace bcm
