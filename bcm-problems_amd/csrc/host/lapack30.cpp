// Symmetric eigen-decomposition and singular value decomposition of small
// dense blocks, restating the LAPACK 3.0 routines the reference's IC2 and
// ILU2 call (lapackd.cpp:865-905 -> lapackdm.f -> lapacksm/dsyev.f,
// dgesvd.f): DSYEV with JOBZ = 'V', UPLO = 'U' and DGESVD with JOBU =
// JOBVT = 'A' on a square matrix; and ZGESVD with JOBU = JOBVT = 'A'
// (lapacksm/zgesvd.f), which the complex ILU2 calls (ilu2cs.cpp:132, 725).
//
// Factor values depend on every rounding of these decompositions, so each
// routine follows its published algorithm operation by operation: the same
// loop orders, the same grouping of sums and products (the reference BLAS
// DDOT's five-term unrolling, DNRM2's scaled sum of squares, DGEMV/DGER's
// column sweeps), the same machine constants (IEEE double: eps = 2^-53,
// safe minimum = 2^-1022). Host code built with -ffp-contract=off, so no
// product is fused into an add.
//
// Blocking: the reference's ILAENV_SM (lapacksm/ilaenv.f) assigns a local
// variable instead of its result for block sizes (ISPEC 1, 2), so the block
// size DSYTRD, DGEBRD and DORGQR/QL/LQ receive is whatever the compiled
// function leaves in its result register (68 in the flang build of the
// reference used as the oracle). For any value of it, the crossover points it
// does return (ISPEC 3: 1 for DSYTRD, 128 for the others) and the workspace
// the factorizations pass (10 n, ich2rs.cpp:56, ilu2rs.cpp:83) make every one
// of those routines take its unblocked path: either the block size is not
// below n, or the workspace holds fewer than that many panels and the block
// size falls under the (equally undefined) minimum. So the call trees are
// DSYTD2, DORG2L, DSTEQR and DGEBD2, DORG2R, DORGL2, DBDSQR for every n.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bcm_host.hpp"

namespace bcm {
namespace lapack30 {

namespace {

constexpr double kEps = 0x1p-53;       // DLAMCH('E'): relative machine precision (rounding)
constexpr double kPrec = 0x1p-52;      // DLAMCH('P') = eps * base
constexpr double kSafmin = 0x1p-1022;  // DLAMCH('S')
// DLARTG's scaling thresholds: base ** int(log(safmin / eps) / log(base) / 2) = 2^-484
constexpr double kSafmn2 = 0x1p-484;
constexpr double kSafmx2 = 0x1p+484;

inline double fsign(double a, double b) { return std::copysign(std::fabs(a), b); }

// ---- reference BLAS ----

// DNRM2: scaled sum of squares, one element at a time.
double dnrm2(int n, const double* x, int incx) {
  if (n < 1 || incx < 1) return 0.0;
  if (n == 1) return std::fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; ++i) {
    const double xi = x[(size_t)i * incx];
    if (xi != 0.0) {
      const double absxi = std::fabs(xi);
      if (scale < absxi) {
        const double q = scale / absxi;
        ssq = 1.0 + ssq * (q * q);
        scale = absxi;
      } else {
        const double q = absxi / scale;
        ssq = ssq + q * q;
      }
    }
  }
  return scale * std::sqrt(ssq);
}

// DDOT with unit strides: the n mod 5 leading terms one at a time, then five
// products added to the running sum left to right.
double ddot(int n, const double* x, const double* y) {
  double t = 0.0;
  if (n <= 0) return t;
  const int m = n % 5;
  for (int i = 0; i < m; ++i) t = t + x[i] * y[i];
  for (int i = m; i < n; i += 5)
    t = t + x[i] * y[i] + x[i + 1] * y[i + 1] + x[i + 2] * y[i + 2] + x[i + 3] * y[i + 3] + x[i + 4] * y[i + 4];
  return t;
}

void daxpy(int n, double a, const double* x, double* y) {
  if (n <= 0 || a == 0.0) return;
  for (int i = 0; i < n; ++i) y[i] = y[i] + a * x[i];
}

void dscal(int n, double a, double* x, int incx) {
  if (n <= 0 || incx <= 0) return;
  for (int i = 0; i < n; ++i) x[(size_t)i * incx] = a * x[(size_t)i * incx];
}

void dswap(int n, double* x, int incx, double* y, int incy) {
  for (int i = 0; i < n; ++i) std::swap(x[(size_t)i * incx], y[(size_t)i * incy]);
}

void drot(int n, double* x, int incx, double* y, int incy, double c, double s) {
  for (int i = 0; i < n; ++i) {
    double& xi = x[(size_t)i * incx];
    double& yi = y[(size_t)i * incy];
    const double t = c * xi + s * yi;
    yi = c * yi - s * xi;
    xi = t;
  }
}

// y = A^T x (DGEMV 'T' with alpha = 1, beta = 0): one column sum at a time.
void gemv_t(int m, int n, const double* a, int lda, const double* x, int incx, double* y) {
  if (m == 0 || n == 0) return;
  for (int j = 0; j < n; ++j) y[j] = 0.0;
  for (int j = 0; j < n; ++j) {
    double t = 0.0;
    for (int i = 0; i < m; ++i) t = t + a[i + (size_t)j * lda] * x[(size_t)i * incx];
    y[j] = y[j] + 1.0 * t;
  }
}

// y = A x (DGEMV 'N' with alpha = 1, beta = 0): column after column, columns
// whose x entry is zero skipped.
void gemv_n(int m, int n, const double* a, int lda, const double* x, int incx, double* y) {
  if (m == 0 || n == 0) return;
  for (int i = 0; i < m; ++i) y[i] = 0.0;
  for (int j = 0; j < n; ++j) {
    const double xj = x[(size_t)j * incx];
    if (xj != 0.0) {
      const double t = 1.0 * xj;
      for (int i = 0; i < m; ++i) y[i] = y[i] + t * a[i + (size_t)j * lda];
    }
  }
}

// A += alpha x y^T (DGER): columns whose y entry is zero skipped.
void dger(int m, int n, double alpha, const double* x, int incx, const double* y, int incy, double* a, int lda) {
  if (m == 0 || n == 0 || alpha == 0.0) return;
  for (int j = 0; j < n; ++j) {
    const double yj = y[(size_t)j * incy];
    if (yj != 0.0) {
      const double t = alpha * yj;
      for (int i = 0; i < m; ++i) a[i + (size_t)j * lda] = a[i + (size_t)j * lda] + x[(size_t)i * incx] * t;
    }
  }
}

// y = alpha A x with A symmetric, upper triangle stored (DSYMV 'U', beta = 0).
void symv_u(int n, double alpha, const double* a, int lda, const double* x, double* y) {
  if (n == 0) return;
  for (int i = 0; i < n; ++i) y[i] = 0.0;
  if (alpha == 0.0) return;
  for (int j = 0; j < n; ++j) {
    const double t1 = alpha * x[j];
    double t2 = 0.0;
    for (int i = 0; i < j; ++i) {
      y[i] = y[i] + t1 * a[i + (size_t)j * lda];
      t2 = t2 + a[i + (size_t)j * lda] * x[i];
    }
    y[j] = y[j] + t1 * a[j + (size_t)j * lda] + alpha * t2;
  }
}

// A += alpha (x y^T + y x^T), upper triangle (DSYR2 'U').
void syr2_u(int n, double alpha, const double* x, const double* y, double* a, int lda) {
  if (n == 0 || alpha == 0.0) return;
  for (int j = 0; j < n; ++j) {
    if (x[j] != 0.0 || y[j] != 0.0) {
      const double t1 = alpha * y[j], t2 = alpha * x[j];
      for (int i = 0; i <= j; ++i)
        a[i + (size_t)j * lda] = a[i + (size_t)j * lda] + x[i] * t1 + y[i] * t2;
    }
  }
}

// ---- LAPACK auxiliaries ----

double dlapy2(double x, double y) {
  const double xa = std::fabs(x), ya = std::fabs(y);
  const double w = std::max(xa, ya), z = std::min(xa, ya);
  if (z == 0.0) return w;
  const double q = z / w;
  return w * std::sqrt(1.0 + q * q);
}

// Elementary reflector H = I - tau v v^T with H (alpha, x) = (beta, 0).
void dlarfg(int n, double& alpha, double* x, int incx, double& tau) {
  if (n <= 1) {
    tau = 0.0;
    return;
  }
  double xnorm = dnrm2(n - 1, x, incx);
  if (xnorm == 0.0) {
    tau = 0.0;
    return;
  }
  double beta = -fsign(dlapy2(alpha, xnorm), alpha);
  const double safmin = kSafmin / kEps;
  if (std::fabs(beta) < safmin) {
    const double rsafmn = 1.0 / safmin;
    int knt = 0;
    do {
      ++knt;
      dscal(n - 1, rsafmn, x, incx);
      beta = beta * rsafmn;
      alpha = alpha * rsafmn;
    } while (std::fabs(beta) < safmin);
    xnorm = dnrm2(n - 1, x, incx);
    beta = -fsign(dlapy2(alpha, xnorm), alpha);
    tau = (beta - alpha) / beta;
    dscal(n - 1, 1.0 / (alpha - beta), x, incx);
    alpha = beta;
    for (int j = 0; j < knt; ++j) alpha = alpha * safmin;
  } else {
    tau = (beta - alpha) / beta;
    dscal(n - 1, 1.0 / (alpha - beta), x, incx);
    alpha = beta;
  }
}

// Applies H = I - tau v v^T from the left (C = H C) or the right (C = C H).
void dlarf(bool left, int m, int n, const double* v, int incv, double tau, double* c, int ldc, double* work) {
  if (tau == 0.0) return;
  if (left) {
    gemv_t(m, n, c, ldc, v, incv, work);
    dger(m, n, -tau, v, incv, work, 1, c, ldc);
  } else {
    gemv_n(m, n, c, ldc, v, incv, work);
    dger(m, n, -tau, work, 1, v, incv, c, ldc);
  }
}

// Plane rotation with r = sqrt(f^2 + g^2) and the sign convention of LAPACK 3.0.
void dlartg(double f, double g, double& cs, double& sn, double& r) {
  if (g == 0.0) {
    cs = 1.0;
    sn = 0.0;
    r = f;
    return;
  }
  if (f == 0.0) {
    cs = 0.0;
    sn = 1.0;
    r = g;
    return;
  }
  double f1 = f, g1 = g;
  double scale = std::max(std::fabs(f1), std::fabs(g1));
  if (scale >= kSafmx2) {
    int count = 0;
    do {
      ++count;
      f1 = f1 * kSafmn2;
      g1 = g1 * kSafmn2;
      scale = std::max(std::fabs(f1), std::fabs(g1));
    } while (scale >= kSafmx2);
    r = std::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * kSafmx2;
  } else if (scale <= kSafmn2) {
    int count = 0;
    do {
      ++count;
      f1 = f1 * kSafmx2;
      g1 = g1 * kSafmx2;
      scale = std::max(std::fabs(f1), std::fabs(g1));
    } while (scale <= kSafmn2);
    r = std::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * kSafmn2;
  } else {
    r = std::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
  }
  if (std::fabs(f) > std::fabs(g) && cs < 0.0) {
    cs = -cs;
    sn = -sn;
    r = -r;
  }
}

// Sequence of plane rotations in planes (j, j+1) (pivot 'V'), from the left
// (rows of a, m x n) or the right (columns), forward or backward.
void dlasr(bool left, bool forward, int m, int n, const double* c, const double* s, double* a, int lda) {
  if (m <= 0 || n <= 0) return;
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  if (left) {
    auto apply = [&](int j) {
      const double ct = c[j], st = s[j];
      if (ct != 1.0 || st != 0.0)
        for (int i = 0; i < n; ++i) {
          const double t = A(j + 1, i);
          A(j + 1, i) = ct * t - st * A(j, i);
          A(j, i) = st * t + ct * A(j, i);
        }
    };
    if (forward)
      for (int j = 0; j < m - 1; ++j) apply(j);
    else
      for (int j = m - 2; j >= 0; --j) apply(j);
  } else {
    auto apply = [&](int j) {
      const double ct = c[j], st = s[j];
      if (ct != 1.0 || st != 0.0)
        for (int i = 0; i < m; ++i) {
          const double t = A(i, j + 1);
          A(i, j + 1) = ct * t - st * A(i, j);
          A(i, j) = st * t + ct * A(i, j);
        }
    };
    if (forward)
      for (int j = 0; j < n - 1; ++j) apply(j);
    else
      for (int j = n - 2; j >= 0; --j) apply(j);
  }
}

// A = A * (cto / cfrom) without over/underflow; type 0 full ('G'), 2 upper triangle ('U').
void dlascl(int type, double cfrom, double cto, int m, int n, double* a, int lda) {
  const double smlnum = kSafmin, bignum = 1.0 / smlnum;
  double cfromc = cfrom, ctoc = cto;
  bool done = false;
  while (!done) {
    const double cfrom1 = cfromc * smlnum;
    const double cto1 = ctoc / bignum;
    double mul;
    if (std::fabs(cfrom1) > std::fabs(ctoc) && ctoc != 0.0) {
      mul = smlnum;
      cfromc = cfrom1;
    } else if (std::fabs(cto1) > std::fabs(cfromc)) {
      mul = bignum;
      ctoc = cto1;
    } else {
      mul = ctoc / cfromc;
      done = true;
    }
    for (int j = 0; j < n; ++j) {
      const int iend = type == 2 ? std::min(j + 1, m) : m;
      for (int i = 0; i < iend; ++i) a[i + (size_t)j * lda] = a[i + (size_t)j * lda] * mul;
    }
  }
}

// Eigenvalues and eigenvector rotation of [[a, b], [b, c]].
void dlaev2(double a, double b, double c, double& rt1, double& rt2, double& cs1, double& sn1) {
  const double sm = a + c, df = a - c, adf = std::fabs(df), tb = b + b, ab = std::fabs(tb);
  double acmx, acmn;
  if (std::fabs(a) > std::fabs(c)) {
    acmx = a;
    acmn = c;
  } else {
    acmx = c;
    acmn = a;
  }
  double rt;
  if (adf > ab) {
    const double q = ab / adf;
    rt = adf * std::sqrt(1.0 + q * q);
  } else if (adf < ab) {
    const double q = adf / ab;
    rt = ab * std::sqrt(1.0 + q * q);
  } else {
    rt = ab * std::sqrt(2.0);
  }
  int sgn1;
  if (sm < 0.0) {
    rt1 = 0.5 * (sm - rt);
    sgn1 = -1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else if (sm > 0.0) {
    rt1 = 0.5 * (sm + rt);
    sgn1 = 1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else {
    rt1 = 0.5 * rt;
    rt2 = -0.5 * rt;
    sgn1 = 1;
  }
  double cs;
  int sgn2;
  if (df >= 0.0) {
    cs = df + rt;
    sgn2 = 1;
  } else {
    cs = df - rt;
    sgn2 = -1;
  }
  const double acs = std::fabs(cs);
  if (acs > ab) {
    const double ct = -tb / cs;
    sn1 = 1.0 / std::sqrt(1.0 + ct * ct);
    cs1 = ct * sn1;
  } else if (ab == 0.0) {
    cs1 = 1.0;
    sn1 = 0.0;
  } else {
    const double tn = -cs / tb;
    cs1 = 1.0 / std::sqrt(1.0 + tn * tn);
    sn1 = tn * cs1;
  }
  if (sgn1 == sgn2) {
    const double tn = cs1;
    cs1 = -sn1;
    sn1 = tn;
  }
}

// Singular values of [[f, g], [0, h]].
void dlas2(double f, double g, double h, double& ssmin, double& ssmax) {
  const double fa = std::fabs(f), ga = std::fabs(g), ha = std::fabs(h);
  const double fhmn = std::min(fa, ha), fhmx = std::max(fa, ha);
  if (fhmn == 0.0) {
    ssmin = 0.0;
    if (fhmx == 0.0) {
      ssmax = ga;
    } else {
      const double q = std::min(fhmx, ga) / std::max(fhmx, ga);
      ssmax = std::max(fhmx, ga) * std::sqrt(1.0 + q * q);
    }
  } else if (ga < fhmx) {
    const double as = 1.0 + fhmn / fhmx;
    const double at = (fhmx - fhmn) / fhmx;
    const double q = ga / fhmx;
    const double au = q * q;
    const double c = 2.0 / (std::sqrt(as * as + au) + std::sqrt(at * at + au));
    ssmin = fhmn * c;
    ssmax = fhmx / c;
  } else {
    const double au = fhmx / ga;
    if (au == 0.0) {
      ssmin = (fhmn * fhmx) / ga;
      ssmax = ga;
    } else {
      const double as = 1.0 + fhmn / fhmx;
      const double at = (fhmx - fhmn) / fhmx;
      const double p = as * au, q = at * au;
      const double c = 1.0 / (std::sqrt(1.0 + p * p) + std::sqrt(1.0 + q * q));
      ssmin = (fhmn * c) * au;
      ssmin = ssmin + ssmin;
      ssmax = ga / (c + c);
    }
  }
}

// SVD of [[f, g], [0, h]] with its left and right rotations.
void dlasv2(double f, double g, double h, double& ssmin, double& ssmax, double& snr, double& csr, double& snl,
            double& csl) {
  double ft = f, fa = std::fabs(ft), ht = h, ha = std::fabs(h);
  int pmax = 1;
  const bool swap = ha > fa;
  if (swap) {
    pmax = 3;
    std::swap(ft, ht);
    std::swap(fa, ha);
  }
  const double gt = g, ga = std::fabs(gt);
  double clt = 0.0, crt = 0.0, slt = 0.0, srt = 0.0;
  if (ga == 0.0) {
    ssmin = ha;
    ssmax = fa;
    clt = 1.0;
    crt = 1.0;
    slt = 0.0;
    srt = 0.0;
  } else {
    bool gasmal = true;
    if (ga > fa) {
      pmax = 2;
      if ((fa / ga) < kEps) {
        gasmal = false;
        ssmax = ga;
        if (ha > 1.0) ssmin = fa / (ga / ha);
        else ssmin = (fa / ga) * ha;
        clt = 1.0;
        slt = ht / gt;
        srt = 1.0;
        crt = ft / gt;
      }
    }
    if (gasmal) {
      const double d = fa - ha;
      double l = d == fa ? 1.0 : d / fa;
      const double m = gt / ft;
      double t = 2.0 - l;
      const double mm = m * m, tt = t * t;
      const double s = std::sqrt(tt + mm);
      const double r = l == 0.0 ? std::fabs(m) : std::sqrt(l * l + mm);
      const double a = 0.5 * (s + r);
      ssmin = ha / a;
      ssmax = fa * a;
      if (mm == 0.0) {
        if (l == 0.0) t = fsign(2.0, ft) * fsign(1.0, gt);
        else t = gt / fsign(d, ft) + m / t;
      } else {
        t = (m / (s + t) + m / (r + l)) * (1.0 + a);
      }
      l = std::sqrt(t * t + 4.0);
      crt = 2.0 / l;
      srt = t / l;
      clt = (crt + srt * m) / a;
      slt = (ht / ft) * srt / a;
    }
  }
  if (swap) {
    csl = srt;
    snl = crt;
    csr = slt;
    snr = clt;
  } else {
    csl = clt;
    snl = slt;
    csr = crt;
    snr = srt;
  }
  double tsign = 1.0;
  if (pmax == 1) tsign = fsign(1.0, csr) * fsign(1.0, csl) * fsign(1.0, f);
  if (pmax == 2) tsign = fsign(1.0, snr) * fsign(1.0, csl) * fsign(1.0, g);
  if (pmax == 3) tsign = fsign(1.0, snr) * fsign(1.0, snl) * fsign(1.0, h);
  ssmax = fsign(ssmax, tsign);
  ssmin = fsign(ssmin, tsign * fsign(1.0, f) * fsign(1.0, h));
}

// ---- DSYEV call tree ----

// Tridiagonal reduction Q^T A Q = T of the upper triangle (DSYTD2 'U').
void dsytd2_u(int n, double* a, int lda, double* d, double* e, double* tau) {
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  for (int i = n - 2; i >= 0; --i) {  // Fortran I = N-1 .. 1 is i + 1
    double taui;
    dlarfg(i + 1, A(i, i + 1), &A(0, i + 1), 1, taui);
    e[i] = A(i, i + 1);
    if (taui != 0.0) {
      A(i, i + 1) = 1.0;
      symv_u(i + 1, taui, a, lda, &A(0, i + 1), tau);
      const double alpha = -(0.5 * taui * ddot(i + 1, tau, &A(0, i + 1)));
      daxpy(i + 1, alpha, &A(0, i + 1), tau);
      syr2_u(i + 1, -1.0, &A(0, i + 1), tau, a, lda);
      A(i, i + 1) = e[i];
    }
    d[i + 1] = A(i + 1, i + 1);
    tau[i] = taui;
  }
  d[0] = A(0, 0);
}

// Generates the m x n matrix Q with orthonormal columns from the last n
// columns of a product of k reflectors (DORG2L).
void dorg2l(int m, int n, int k, double* a, int lda, const double* tau, double* work) {
  if (n <= 0) return;
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  for (int j = 0; j < n - k; ++j) {
    for (int l = 0; l < m; ++l) A(l, j) = 0.0;
    A(m - n + j, j) = 1.0;
  }
  for (int i = 0; i < k; ++i) {
    const int ii = n - k + i;  // 0-based column; Fortran II = ii + 1
    A(m - n + ii, ii) = 1.0;
    dlarf(true, m - n + ii + 1, ii, &A(0, ii), 1, tau[i], a, lda, work);
    dscal(m - n + ii, -tau[i], &A(0, ii), 1);
    A(m - n + ii, ii) = 1.0 - tau[i];
    for (int l = m - n + ii + 1; l < m; ++l) A(l, ii) = 0.0;
  }
}

// Q of DSYTD2 'U' (DORGTR 'U' -> DORGQL -> DORG2L for n - 1 <= 128).
void dorgtr_u(int n, double* a, int lda, const double* tau, double* work) {
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  for (int j = 0; j < n - 1; ++j) {
    for (int i = 0; i < j; ++i) A(i, j) = A(i, j + 1);
    A(n - 1, j) = 0.0;
  }
  for (int i = 0; i < n - 1; ++i) A(i, n - 1) = 0.0;
  A(n - 1, n - 1) = 1.0;
  dorg2l(n - 1, n - 1, n - 1, a, lda, tau, work);
}

double dlanst_i(int n, const double* d, const double* e) {
  if (n <= 0) return 0.0;
  if (n == 1) return std::fabs(d[0]);
  double anorm = std::max(std::fabs(d[0]) + std::fabs(e[0]), std::fabs(e[n - 2]) + std::fabs(d[n - 1]));
  for (int i = 1; i < n - 1; ++i) anorm = std::max(anorm, std::fabs(d[i]) + std::fabs(e[i]) + std::fabs(e[i - 1]));
  return anorm;
}

// Implicit QL/QR on the symmetric tridiagonal (d, e), accumulating the
// rotations into z (DSTEQR with COMPZ = 'V'); eigenvalues sorted ascending
// by selection sort with their columns of z. Returns INFO.
int dsteqr_v(int n, double* d, double* e, double* z, int ldz, double* work) {
  if (n <= 1) return 0;
  const int maxit = 30;
  const double eps = kEps, eps2 = eps * eps, safmin = kSafmin, safmax = 1.0 / safmin;
  const double ssfmax = std::sqrt(safmax) / 3.0, ssfmin = std::sqrt(safmin) / eps2;
  // 1-based views, as the algorithm is written
  auto D = [&](int i) -> double& { return d[i - 1]; };
  auto E = [&](int i) -> double& { return e[i - 1]; };
  auto W = [&](int i) -> double& { return work[i - 1]; };
  auto Zc = [&](int j) { return z + (size_t)(j - 1) * ldz; };
  const int nmaxit = n * maxit;
  int jtot = 0, l1 = 1, info = 0;
  const int nm1 = n - 1;
  while (true) {
    // label 10
    if (l1 > n) break;  // -> 160
    if (l1 > 1) E(l1 - 1) = 0.0;
    int m = n;
    if (l1 <= nm1) {
      for (m = l1; m <= nm1; ++m) {
        const double tst = std::fabs(E(m));
        if (tst == 0.0) break;
        if (tst <= (std::sqrt(std::fabs(D(m))) * std::sqrt(std::fabs(D(m + 1)))) * eps) {
          E(m) = 0.0;
          break;
        }
      }
      if (m > nm1) m = n;
    }
    int l = l1;
    const int lsv = l;
    int lend = m;
    const int lendsv = lend;
    l1 = m + 1;
    if (lend == l) continue;
    const double anorm = dlanst_i(lend - l + 1, &D(l), &E(l));
    int iscale = 0;
    if (anorm == 0.0) continue;
    if (anorm > ssfmax) {
      iscale = 1;
      dlascl(0, anorm, ssfmax, lend - l + 1, 1, &D(l), n);
      dlascl(0, anorm, ssfmax, lend - l, 1, &E(l), n);
    } else if (anorm < ssfmin) {
      iscale = 2;
      dlascl(0, anorm, ssfmin, lend - l + 1, 1, &D(l), n);
      dlascl(0, anorm, ssfmin, lend - l, 1, &E(l), n);
    }
    if (std::fabs(D(lend)) < std::fabs(D(l))) {
      lend = lsv;
      l = lendsv;
    }
    if (lend > l) {
      // QL iteration
      while (true) {  // label 40
        int mm;
        if (l != lend) {
          const int lendm1 = lend - 1;
          for (mm = l; mm <= lendm1; ++mm) {
            const double tst = std::fabs(E(mm)) * std::fabs(E(mm));
            if (tst <= (eps2 * std::fabs(D(mm))) * std::fabs(D(mm + 1)) + safmin) break;
          }
          if (mm > lendm1) mm = lend;
        } else {
          mm = lend;
        }
        m = mm;
        if (m < lend) E(m) = 0.0;
        double p = D(l);
        if (m == l) {  // label 80
          D(l) = p;
          l = l + 1;
          if (l <= lend) continue;
          break;  // -> 140
        }
        if (m == l + 1) {
          double rt1, rt2, c, s;
          dlaev2(D(l), E(l), D(l + 1), rt1, rt2, c, s);
          W(l) = c;
          W(n - 1 + l) = s;
          dlasr(false, false, n, 2, &W(l), &W(n - 1 + l), Zc(l), ldz);
          D(l) = rt1;
          D(l + 1) = rt2;
          E(l) = 0.0;
          l = l + 2;
          if (l <= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        jtot = jtot + 1;
        double g = (D(l + 1) - p) / (2.0 * E(l));
        double r = dlapy2(g, 1.0);
        g = D(m) - p + (E(l) / (g + fsign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        const int mm1 = m - 1;
        for (int i = mm1; i >= l; --i) {
          const double f = s * E(i);
          const double b = c * E(i);
          dlartg(g, f, c, s, r);
          if (i != m - 1) E(i + 1) = r;
          g = D(i + 1) - p;
          r = (D(i) - g) * s + 2.0 * c * b;
          p = s * r;
          D(i + 1) = g + p;
          g = c * r - b;
          W(i) = c;
          W(n - 1 + i) = -s;
        }
        dlasr(false, false, n, m - l + 1, &W(l), &W(n - 1 + l), Zc(l), ldz);
        D(l) = D(l) - p;
        E(l) = g;
      }
    } else {
      // QR iteration
      while (true) {  // label 90
        int mm;
        if (l != lend) {
          const int lendp1 = lend + 1;
          for (mm = l; mm >= lendp1; --mm) {
            const double tst = std::fabs(E(mm - 1)) * std::fabs(E(mm - 1));
            if (tst <= (eps2 * std::fabs(D(mm))) * std::fabs(D(mm - 1)) + safmin) break;
          }
          if (mm < lendp1) mm = lend;
        } else {
          mm = lend;
        }
        m = mm;
        if (m > lend) E(m - 1) = 0.0;
        double p = D(l);
        if (m == l) {  // label 130
          D(l) = p;
          l = l - 1;
          if (l >= lend) continue;
          break;
        }
        if (m == l - 1) {
          double rt1, rt2, c, s;
          dlaev2(D(l - 1), E(l - 1), D(l), rt1, rt2, c, s);
          W(m) = c;
          W(n - 1 + m) = s;
          dlasr(false, true, n, 2, &W(m), &W(n - 1 + m), Zc(l - 1), ldz);
          D(l - 1) = rt1;
          D(l) = rt2;
          E(l - 1) = 0.0;
          l = l - 2;
          if (l >= lend) continue;
          break;
        }
        if (jtot == nmaxit) break;
        jtot = jtot + 1;
        double g = (D(l - 1) - p) / (2.0 * E(l - 1));
        double r = dlapy2(g, 1.0);
        g = D(m) - p + (E(l - 1) / (g + fsign(r, g)));
        double s = 1.0, c = 1.0;
        p = 0.0;
        const int lm1 = l - 1;
        for (int i = m; i <= lm1; ++i) {
          const double f = s * E(i);
          const double b = c * E(i);
          dlartg(g, f, c, s, r);
          if (i != m) E(i - 1) = r;
          g = D(i) - p;
          r = (D(i + 1) - g) * s + 2.0 * c * b;
          p = s * r;
          D(i) = g + p;
          g = c * r - b;
          W(i) = c;
          W(n - 1 + i) = s;
        }
        dlasr(false, true, n, l - m + 1, &W(m), &W(n - 1 + m), Zc(m), ldz);
        D(l) = D(l) - p;
        E(lm1) = g;
      }
    }
    // label 140
    if (iscale == 1) {
      dlascl(0, ssfmax, anorm, lendsv - lsv + 1, 1, &D(lsv), n);
      dlascl(0, ssfmax, anorm, lendsv - lsv, 1, &E(lsv), n);
    } else if (iscale == 2) {
      dlascl(0, ssfmin, anorm, lendsv - lsv + 1, 1, &D(lsv), n);
      dlascl(0, ssfmin, anorm, lendsv - lsv, 1, &E(lsv), n);
    }
    if (jtot < nmaxit) continue;
    for (int i = 1; i <= n - 1; ++i)
      if (E(i) != 0.0) ++info;
    return info;
  }
  // label 160: selection sort, ascending
  for (int ii = 2; ii <= n; ++ii) {
    const int i = ii - 1;
    int k = i;
    double p = D(i);
    for (int j = ii; j <= n; ++j)
      if (D(j) < p) {
        k = j;
        p = D(j);
      }
    if (k != i) {
      D(k) = D(i);
      D(i) = p;
      dswap(n, Zc(i), 1, Zc(k), 1);
    }
  }
  return info;
}

// ---- DGESVD call tree (square, JOBU = JOBVT = 'A') ----

// Bidiagonal reduction Q^T A P = B, upper bidiagonal for m >= n (DGEBD2).
void dgebd2(int m, int n, double* a, int lda, double* d, double* e, double* tauq, double* taup, double* work) {
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  for (int i = 0; i < n; ++i) {
    dlarfg(m - i, A(i, i), &A(std::min(i + 1, m - 1), i), 1, tauq[i]);
    d[i] = A(i, i);
    A(i, i) = 1.0;
    dlarf(true, m - i, n - i - 1, &A(i, i), 1, tauq[i], &A(i, i + 1 < n ? i + 1 : i), lda, work);
    A(i, i) = d[i];
    if (i < n - 1) {
      dlarfg(n - i - 1, A(i, i + 1), &A(i, std::min(i + 2, n - 1)), lda, taup[i]);
      e[i] = A(i, i + 1);
      A(i, i + 1) = 1.0;
      dlarf(false, m - i - 1, n - i - 1, &A(i, i + 1), lda, taup[i], &A(i + 1 < m ? i + 1 : i, i + 1), lda, work);
      A(i, i + 1) = e[i];
    } else {
      taup[i] = 0.0;
    }
  }
}

// Q with orthonormal columns from k reflectors stored in the columns (DORG2R).
void dorg2r(int m, int n, int k, double* a, int lda, const double* tau, double* work) {
  if (n <= 0) return;
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  for (int j = k; j < n; ++j) {
    for (int l = 0; l < m; ++l) A(l, j) = 0.0;
    A(j, j) = 1.0;
  }
  for (int i = k - 1; i >= 0; --i) {
    if (i < n - 1) {
      A(i, i) = 1.0;
      dlarf(true, m - i, n - i - 1, &A(i, i), 1, tau[i], &A(i, i + 1), lda, work);
    }
    if (i < m - 1) dscal(m - i - 1, -tau[i], &A(i + 1, i), 1);
    A(i, i) = 1.0 - tau[i];
    for (int l = 0; l < i; ++l) A(l, i) = 0.0;
  }
}

// Q with orthonormal rows from k reflectors stored in the rows (DORGL2).
void dorgl2(int m, int n, int k, double* a, int lda, const double* tau, double* work) {
  if (m <= 0) return;
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  if (k < m) {
    for (int j = 0; j < n; ++j) {
      for (int l = k; l < m; ++l) A(l, j) = 0.0;
      if (j >= k && j < m) A(j, j) = 1.0;
    }
  }
  for (int i = k - 1; i >= 0; --i) {
    if (i < n - 1) {
      if (i < m - 1) {
        A(i, i) = 1.0;
        dlarf(false, m - i - 1, n - i, &A(i, i), lda, tau[i], &A(i + 1, i), lda, work);
      }
      dscal(n - i - 1, -tau[i], &A(i, i + 1), lda);
    }
    A(i, i) = 1.0 - tau[i];
    for (int l = 0; l < i; ++l) A(i, l) = 0.0;
  }
}

// Q of the bidiagonal reduction, m >= k (DORGBR 'Q' -> DORGQR -> DORG2R).
void dorgbr_q(int m, int n, int k, double* a, int lda, const double* tau, double* work) {
  if (m == 0 || n == 0) return;
  dorg2r(m, n, k, a, lda, tau, work);
}

// P^T of the bidiagonal reduction of a square n x n matrix (DORGBR 'P' with
// k >= n: shift the reflectors one row down, then DORGLQ -> DORGL2).
void dorgbr_p_square(int n, double* a, int lda, const double* tau, double* work) {
  if (n == 0) return;
  auto A = [&](int i, int j) -> double& { return a[i + (size_t)j * lda]; };
  A(0, 0) = 1.0;
  for (int i = 1; i < n; ++i) A(i, 0) = 0.0;
  for (int j = 1; j < n; ++j) {
    for (int i = j - 1; i >= 1; --i) A(i, j) = A(i - 1, j);
    A(0, j) = 0.0;
  }
  if (n > 1) dorgl2(n - 1, n - 1, n - 1, &A(1, 1), lda, tau, work);
}

// ---- the vector updates of the bidiagonal QR, on real or complex vectors ----
//
// ZBDSQR is DBDSQR with its rotations, sign flips and swaps applied to complex
// vectors (ZDROT, ZLASR, ZDSCAL, ZSWAP); the bidiagonal itself stays real.

inline void rot_v(int n, double* x, int incx, double* y, int incy, double c, double s) { drot(n, x, incx, y, incy, c, s); }
inline void lasr_v(bool left, bool forward, int m, int n, const double* c, const double* s, double* a, int lda) {
  dlasr(left, forward, m, n, c, s, a, lda);
}
inline void scal_v(int n, double a, double* x, int incx) { dscal(n, a, x, incx); }
inline void swap_v(int n, double* x, int incx, double* y, int incy) { dswap(n, x, incx, y, incy); }

// Complex arithmetic as the compiled Fortran evaluates it: a product of two
// complex values is (ac - bd, ad + bc); a real times a complex value scales
// both parts. The two forms differ at most in the sign of a zero part.
inline Zc zmul(Zc a, Zc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline Zc zadd(Zc a, Zc b) { return {a.re + b.re, a.im + b.im}; }
inline Zc zsub(Zc a, Zc b) { return {a.re - b.re, a.im - b.im}; }
inline Zc zconj(Zc a) { return {a.re, -a.im}; }
inline Zc zneg(Zc a) { return {-a.re, -a.im}; }
inline Zc zrs(double c, Zc a) { return {c * a.re, c * a.im}; }
inline bool znonzero(Zc a) { return a.re != 0.0 || a.im != 0.0; }
constexpr Zc kZero{0.0, 0.0};
constexpr Zc kOne{1.0, 0.0};

// ZDROT: x, y <- c x + s y, c y - s x.
void rot_v(int n, Zc* x, int incx, Zc* y, int incy, double c, double s) {
  for (int i = 0; i < n; ++i) {
    Zc& xi = x[(size_t)i * incx];
    Zc& yi = y[(size_t)i * incy];
    const Zc t = zadd(zrs(c, xi), zrs(s, yi));
    yi = zsub(zrs(c, yi), zrs(s, xi));
    xi = t;
  }
}

// ZLASR: dlasr on complex entries.
void lasr_v(bool left, bool forward, int m, int n, const double* c, const double* s, Zc* a, int lda) {
  if (m <= 0 || n <= 0) return;
  auto A = [&](int i, int j) -> Zc& { return a[i + (size_t)j * lda]; };
  auto rotate = [&](Zc& lo, Zc& hi, double ct, double st) {
    const Zc t = hi;
    hi = zsub(zrs(ct, t), zrs(st, lo));
    lo = zadd(zrs(st, t), zrs(ct, lo));
  };
  if (left) {
    auto apply = [&](int j) {
      const double ct = c[j], st = s[j];
      if (ct != 1.0 || st != 0.0)
        for (int i = 0; i < n; ++i) rotate(A(j, i), A(j + 1, i), ct, st);
    };
    if (forward)
      for (int j = 0; j < m - 1; ++j) apply(j);
    else
      for (int j = m - 2; j >= 0; --j) apply(j);
  } else {
    auto apply = [&](int j) {
      const double ct = c[j], st = s[j];
      if (ct != 1.0 || st != 0.0)
        for (int i = 0; i < m; ++i) rotate(A(i, j), A(i, j + 1), ct, st);
    };
    if (forward)
      for (int j = 0; j < n - 1; ++j) apply(j);
    else
      for (int j = n - 2; j >= 0; --j) apply(j);
  }
}

// ZDSCAL: x <- dcmplx(a, 0) x, a complex product.
void scal_v(int n, double a, Zc* x, int incx) {
  if (n <= 0) return;
  for (int i = 0; i < n; ++i) x[(size_t)i * incx] = zmul(Zc{a, 0.0}, x[(size_t)i * incx]);
}

void swap_v(int n, Zc* x, int incx, Zc* y, int incy) {
  for (int i = 0; i < n; ++i) std::swap(x[(size_t)i * incx], y[(size_t)i * incy]);
}

// SVD of an upper bidiagonal (d, e) by implicit zero-shift / shifted QR,
// rotating the rows of vt (n x ncvt) and the columns of u (nru x n) (DBDSQR
// 'U' with NCC = 0 for real T, ZBDSQR for complex T). Returns INFO.
template <class T>
int bdsqr_u(int n, int ncvt, int nru, double* d, double* e, T* vt, int ldvt, T* u, int ldu,
             double* work) {
  if (n == 0) return 0;
  auto D = [&](int i) -> double& { return d[i - 1]; };
  auto E = [&](int i) -> double& { return e[i - 1]; };
  auto W = [&](int i) -> double& { return work[i - 1]; };
  auto VT = [&](int i, int j) -> T& { return vt[(i - 1) + (size_t)(j - 1) * ldvt]; };
  auto U = [&](int i, int j) -> T& { return u[(i - 1) + (size_t)(j - 1) * ldu]; };
  const double hndrth = 0.01, ten = 10.0, hndrd = 100.0, meigth = -0.125;
  const int maxitr = 6;
  int info = 0;
  if (n > 1) {
    const int nm1 = n - 1, nm12 = nm1 + nm1, nm13 = nm12 + nm1;
    int idir = 0;
    const double eps = kEps, unfl = kSafmin;
    const double tolmul = std::max(ten, std::min(hndrd, std::pow(eps, meigth)));
    const double tol = tolmul * eps;
    double smax = 0.0;
    for (int i = 1; i <= n; ++i) smax = std::max(smax, std::fabs(D(i)));
    for (int i = 1; i <= n - 1; ++i) smax = std::max(smax, std::fabs(E(i)));
    double sminl = 0.0, thresh;
    if (tol >= 0.0) {
      double sminoa = std::fabs(D(1));
      if (sminoa != 0.0) {
        double mu = sminoa;
        for (int i = 2; i <= n; ++i) {
          mu = std::fabs(D(i)) * (mu / (mu + std::fabs(E(i - 1))));
          sminoa = std::min(sminoa, mu);
          if (sminoa == 0.0) break;
        }
      }
      sminoa = sminoa / std::sqrt((double)n);
      thresh = std::max(tol * sminoa, (double)(maxitr * n * n) * unfl);
    } else {
      thresh = std::max(std::fabs(tol) * smax, (double)(maxitr * n * n) * unfl);
    }
    const int maxit = maxitr * n * n;
    int iter = 0, oldll = -1, oldm = -1, m = n;
    while (true) {  // label 60
      if (m <= 1) break;  // -> 160
      if (iter > maxit) {  // label 200
        info = 0;
        for (int i = 1; i <= n - 1; ++i)
          if (E(i) != 0.0) ++info;
        return info;
      }
      if (tol < 0.0 && std::fabs(D(m)) <= thresh) D(m) = 0.0;
      smax = std::fabs(D(m));
      double smin = smax;
      int ll = 0;
      bool split = false;
      for (int lll = 1; lll <= m - 1; ++lll) {
        ll = m - lll;
        const double abss = std::fabs(D(ll));
        const double abse = std::fabs(E(ll));
        if (tol < 0.0 && abss <= thresh) D(ll) = 0.0;
        if (abse <= thresh) {
          split = true;
          break;
        }
        smin = std::min(smin, abss);
        smax = std::max(smax, std::max(abss, abse));
      }
      if (split) {  // label 80
        E(ll) = 0.0;
        if (ll == m - 1) {
          m = m - 1;
          continue;
        }
      } else {
        ll = 0;
      }
      // label 90
      ll = ll + 1;
      if (ll == m - 1) {
        double sigmn, sigmx, sinr, cosr, sinl, cosl;
        dlasv2(D(m - 1), E(m - 1), D(m), sigmn, sigmx, sinr, cosr, sinl, cosl);
        D(m - 1) = sigmx;
        E(m - 1) = 0.0;
        D(m) = sigmn;
        if (ncvt > 0) rot_v(ncvt, &VT(m - 1, 1), ldvt, &VT(m, 1), ldvt, cosr, sinr);
        if (nru > 0) rot_v(nru, &U(1, m - 1), 1, &U(1, m), 1, cosl, sinl);
        m = m - 2;
        continue;
      }
      if (ll > oldm || m < oldll) {
        if (std::fabs(D(ll)) >= std::fabs(D(m))) idir = 1;
        else idir = 2;
      }
      bool restart = false;
      if (idir == 1) {
        if (std::fabs(E(m - 1)) <= std::fabs(tol) * std::fabs(D(m)) ||
            (tol < 0.0 && std::fabs(E(m - 1)) <= thresh)) {
          E(m - 1) = 0.0;
          continue;
        }
        if (tol >= 0.0) {
          double mu = std::fabs(D(ll));
          sminl = mu;
          for (int lll = ll; lll <= m - 1; ++lll) {
            if (std::fabs(E(lll)) <= tol * mu) {
              E(lll) = 0.0;
              restart = true;
              break;
            }
            mu = std::fabs(D(lll + 1)) * (mu / (mu + std::fabs(E(lll))));
            sminl = std::min(sminl, mu);
          }
          if (restart) continue;
        }
      } else {
        if (std::fabs(E(ll)) <= std::fabs(tol) * std::fabs(D(ll)) || (tol < 0.0 && std::fabs(E(ll)) <= thresh)) {
          E(ll) = 0.0;
          continue;
        }
        if (tol >= 0.0) {
          double mu = std::fabs(D(m));
          sminl = mu;
          for (int lll = m - 1; lll >= ll; --lll) {
            if (std::fabs(E(lll)) <= tol * mu) {
              E(lll) = 0.0;
              restart = true;
              break;
            }
            mu = std::fabs(D(lll)) * (mu / (mu + std::fabs(E(lll))));
            sminl = std::min(sminl, mu);
          }
          if (restart) continue;
        }
      }
      oldll = ll;
      oldm = m;
      double shift, r;
      if (tol >= 0.0 && (double)n * tol * (sminl / smax) <= std::max(eps, hndrth * tol)) {
        shift = 0.0;
      } else {
        double sll;
        if (idir == 1) {
          sll = std::fabs(D(ll));
          dlas2(D(m - 1), E(m - 1), D(m), shift, r);
        } else {
          sll = std::fabs(D(m));
          dlas2(D(ll), E(ll), D(ll + 1), shift, r);
        }
        if (sll > 0.0) {
          const double q = shift / sll;
          if (q * q < eps) shift = 0.0;
        }
      }
      iter = iter + m - ll;
      if (shift == 0.0) {
        if (idir == 1) {
          double cs = 1.0, oldcs = 1.0, sn = 0.0, oldsn = 0.0;
          for (int i = ll; i <= m - 1; ++i) {
            dlartg(D(i) * cs, E(i), cs, sn, r);
            if (i > ll) E(i - 1) = oldsn * r;
            dlartg(oldcs * r, D(i + 1) * sn, oldcs, oldsn, D(i));
            W(i - ll + 1) = cs;
            W(i - ll + 1 + nm1) = sn;
            W(i - ll + 1 + nm12) = oldcs;
            W(i - ll + 1 + nm13) = oldsn;
          }
          const double h = D(m) * cs;
          D(m) = h * oldcs;
          E(m - 1) = h * oldsn;
          if (ncvt > 0) lasr_v(true, true, m - ll + 1, ncvt, &W(1), &W(n), &VT(ll, 1), ldvt);
          if (nru > 0) lasr_v(false, true, nru, m - ll + 1, &W(nm12 + 1), &W(nm13 + 1), &U(1, ll), ldu);
          if (std::fabs(E(m - 1)) <= thresh) E(m - 1) = 0.0;
        } else {
          double cs = 1.0, oldcs = 1.0, sn = 0.0, oldsn = 0.0;
          for (int i = m; i >= ll + 1; --i) {
            dlartg(D(i) * cs, E(i - 1), cs, sn, r);
            if (i < m) E(i) = oldsn * r;
            dlartg(oldcs * r, D(i - 1) * sn, oldcs, oldsn, D(i));
            W(i - ll) = cs;
            W(i - ll + nm1) = -sn;
            W(i - ll + nm12) = oldcs;
            W(i - ll + nm13) = -oldsn;
          }
          const double h = D(ll) * cs;
          D(ll) = h * oldcs;
          E(ll) = h * oldsn;
          if (ncvt > 0) lasr_v(true, false, m - ll + 1, ncvt, &W(nm12 + 1), &W(nm13 + 1), &VT(ll, 1), ldvt);
          if (nru > 0) lasr_v(false, false, nru, m - ll + 1, &W(1), &W(n), &U(1, ll), ldu);
          if (std::fabs(E(ll)) <= thresh) E(ll) = 0.0;
        }
      } else {
        if (idir == 1) {
          double f = (std::fabs(D(ll)) - shift) * (fsign(1.0, D(ll)) + shift / D(ll));
          double g = E(ll);
          double cosr, sinr, cosl, sinl;
          for (int i = ll; i <= m - 1; ++i) {
            dlartg(f, g, cosr, sinr, r);
            if (i > ll) E(i - 1) = r;
            f = cosr * D(i) + sinr * E(i);
            E(i) = cosr * E(i) - sinr * D(i);
            g = sinr * D(i + 1);
            D(i + 1) = cosr * D(i + 1);
            dlartg(f, g, cosl, sinl, r);
            D(i) = r;
            f = cosl * E(i) + sinl * D(i + 1);
            D(i + 1) = cosl * D(i + 1) - sinl * E(i);
            if (i < m - 1) {
              g = sinl * E(i + 1);
              E(i + 1) = cosl * E(i + 1);
            }
            W(i - ll + 1) = cosr;
            W(i - ll + 1 + nm1) = sinr;
            W(i - ll + 1 + nm12) = cosl;
            W(i - ll + 1 + nm13) = sinl;
          }
          E(m - 1) = f;
          if (ncvt > 0) lasr_v(true, true, m - ll + 1, ncvt, &W(1), &W(n), &VT(ll, 1), ldvt);
          if (nru > 0) lasr_v(false, true, nru, m - ll + 1, &W(nm12 + 1), &W(nm13 + 1), &U(1, ll), ldu);
          if (std::fabs(E(m - 1)) <= thresh) E(m - 1) = 0.0;
        } else {
          double f = (std::fabs(D(m)) - shift) * (fsign(1.0, D(m)) + shift / D(m));
          double g = E(m - 1);
          double cosr, sinr, cosl, sinl;
          for (int i = m; i >= ll + 1; --i) {
            dlartg(f, g, cosr, sinr, r);
            if (i < m) E(i) = r;
            f = cosr * D(i) + sinr * E(i - 1);
            E(i - 1) = cosr * E(i - 1) - sinr * D(i);
            g = sinr * D(i - 1);
            D(i - 1) = cosr * D(i - 1);
            dlartg(f, g, cosl, sinl, r);
            D(i) = r;
            f = cosl * E(i - 1) + sinl * D(i - 1);
            D(i - 1) = cosl * D(i - 1) - sinl * E(i - 1);
            if (i > ll + 1) {
              g = sinl * E(i - 2);
              E(i - 2) = cosl * E(i - 2);
            }
            W(i - ll) = cosr;
            W(i - ll + nm1) = -sinr;
            W(i - ll + nm12) = cosl;
            W(i - ll + nm13) = -sinl;
          }
          E(ll) = f;
          if (std::fabs(E(ll)) <= thresh) E(ll) = 0.0;
          if (ncvt > 0) lasr_v(true, false, m - ll + 1, ncvt, &W(nm12 + 1), &W(nm13 + 1), &VT(ll, 1), ldvt);
          if (nru > 0) lasr_v(false, false, nru, m - ll + 1, &W(1), &W(n), &U(1, ll), ldu);
        }
      }
    }
  }
  // label 160: make the singular values positive, sort them decreasing
  for (int i = 1; i <= n; ++i) {
    if (D(i) < 0.0) {
      D(i) = -D(i);
      if (ncvt > 0) scal_v(ncvt, -1.0, &VT(i, 1), ldvt);
    }
  }
  for (int i = 1; i <= n - 1; ++i) {
    int isub = 1;
    double smin = D(1);
    for (int j = 2; j <= n + 1 - i; ++j)
      if (D(j) <= smin) {
        isub = j;
        smin = D(j);
      }
    if (isub != n + 1 - i) {
      D(isub) = D(n + 1 - i);
      D(n + 1 - i) = smin;
      if (ncvt > 0) swap_v(ncvt, &VT(isub, 1), ldvt, &VT(n + 1 - i, 1), ldvt);
      if (nru > 0) swap_v(nru, &U(1, isub), 1, &U(1, n + 1 - i), 1);
    }
  }
  return info;
}

// ---- ZGESVD call tree (square, JOBU = JOBVT = 'A') ----
//
// Under the same blocking argument as the real tree, ZGEBRD, ZUNGQR and
// ZUNGLQ take their unblocked paths: ZGEBD2, ZUNG2R, ZUNGL2.

// DZNRM2: the scaled sum of squares over the real and the imaginary part of
// each entry in turn.
double dznrm2(int n, const Zc* x, int incx) {
  if (n < 1 || incx < 1) return 0.0;
  double scale = 0.0, ssq = 1.0;
  auto add = [&](double v) {
    if (v != 0.0) {
      const double temp = std::fabs(v);
      if (scale < temp) {
        const double q = scale / temp;
        ssq = 1.0 + ssq * (q * q);
        scale = temp;
      } else {
        const double q = temp / scale;
        ssq = ssq + q * q;
      }
    }
  };
  for (int i = 0; i < n; ++i) {
    add(x[(size_t)i * incx].re);
    add(x[(size_t)i * incx].im);
  }
  return scale * std::sqrt(ssq);
}

// sqrt(x^2 + y^2 + z^2) scaled by the largest magnitude (DLAPY3).
double dlapy3(double x, double y, double z) {
  const double xa = std::fabs(x), ya = std::fabs(y), za = std::fabs(z);
  const double w = std::max(std::max(xa, ya), za);
  if (w == 0.0) return 0.0;
  const double qx = xa / w, qy = ya / w, qz = za / w;
  return w * std::sqrt(qx * qx + qy * qy + qz * qz);
}

// (a + ib) / (c + id) by Smith's scaling (DLADIV; ZLADIV calls it).
void dladiv(double a, double b, double c, double d, double& p, double& q) {
  if (std::fabs(d) < std::fabs(c)) {
    const double e = d / c, f = c + d * e;
    p = (a + b * e) / f;
    q = (b - a * e) / f;
  } else {
    const double e = c / d, f = d + c * e;
    p = (b + a * e) / f;
    q = (-a + b * e) / f;
  }
}

// ZSCAL: x <- a x.
void zscal(int n, Zc a, Zc* x, int incx) {
  for (int i = 0; i < n; ++i) x[(size_t)i * incx] = zmul(a, x[(size_t)i * incx]);
}

// ZLACGV: conjugates n entries.
void zlacgv(int n, Zc* x, int incx) {
  for (int i = 0; i < n; ++i) x[(size_t)i * incx].im = -x[(size_t)i * incx].im;
}

// Elementary reflector H = I - tau v v^H with H^H (alpha, x) = (beta, 0),
// beta real (ZLARFG): unlike the real reflector, it is formed for n = 1 as
// well whenever alpha is not real.
void zlarfg(int n, Zc& alpha, Zc* x, int incx, Zc& tau) {
  if (n <= 0) {
    tau = kZero;
    return;
  }
  double xnorm = dznrm2(n - 1, x, incx);
  double alphr = alpha.re, alphi = alpha.im;
  if (xnorm == 0.0 && alphi == 0.0) {
    tau = kZero;
    return;
  }
  double beta = -fsign(dlapy3(alphr, alphi, xnorm), alphr);
  const double safmin = kSafmin / kEps, rsafmn = 1.0 / safmin;
  int knt = 0;
  if (std::fabs(beta) < safmin) {
    do {
      ++knt;
      scal_v(n - 1, rsafmn, x, incx);
      beta = beta * rsafmn;
      alphi = alphi * rsafmn;
      alphr = alphr * rsafmn;
    } while (std::fabs(beta) < safmin);
    xnorm = dznrm2(n - 1, x, incx);
    alpha = Zc{alphr, alphi};
    beta = -fsign(dlapy3(alphr, alphi, xnorm), alphr);
  }
  tau = Zc{(beta - alphr) / beta, -alphi / beta};
  Zc inv;
  dladiv(1.0, 0.0, alpha.re - beta, alpha.im, inv.re, inv.im);  // 1 / (alpha - beta)
  zscal(n - 1, inv, x, incx);
  alpha = Zc{beta, 0.0};
  for (int j = 0; j < knt; ++j) alpha = zrs(safmin, alpha);
}

// y = A^H x (ZGEMV 'C' with alpha = 1, beta = 0): one column sum at a time.
void zgemv_c(int m, int n, const Zc* a, int lda, const Zc* x, int incx, Zc* y) {
  if (m == 0 || n == 0) return;
  for (int j = 0; j < n; ++j) y[j] = kZero;
  for (int j = 0; j < n; ++j) {
    Zc t = kZero;
    for (int i = 0; i < m; ++i) t = zadd(t, zmul(zconj(a[i + (size_t)j * lda]), x[(size_t)i * incx]));
    y[j] = zadd(y[j], zmul(kOne, t));
  }
}

// y = A x (ZGEMV 'N' with alpha = 1, beta = 0): columns whose x entry is zero skipped.
void zgemv_n(int m, int n, const Zc* a, int lda, const Zc* x, int incx, Zc* y) {
  if (m == 0 || n == 0) return;
  for (int i = 0; i < m; ++i) y[i] = kZero;
  for (int j = 0; j < n; ++j) {
    const Zc xj = x[(size_t)j * incx];
    if (znonzero(xj)) {
      const Zc t = zmul(kOne, xj);
      for (int i = 0; i < m; ++i) y[i] = zadd(y[i], zmul(t, a[i + (size_t)j * lda]));
    }
  }
}

// A += alpha x y^H (ZGERC): columns whose y entry is zero skipped.
void zgerc(int m, int n, Zc alpha, const Zc* x, int incx, const Zc* y, int incy, Zc* a, int lda) {
  if (m == 0 || n == 0 || !znonzero(alpha)) return;
  for (int j = 0; j < n; ++j) {
    const Zc yj = y[(size_t)j * incy];
    if (znonzero(yj)) {
      const Zc t = zmul(alpha, zconj(yj));
      for (int i = 0; i < m; ++i) a[i + (size_t)j * lda] = zadd(a[i + (size_t)j * lda], zmul(x[(size_t)i * incx], t));
    }
  }
}

// Applies H = I - tau v v^H from the left (C = H C) or the right (C = C H) (ZLARF).
void zlarf(bool left, int m, int n, const Zc* v, int incv, Zc tau, Zc* c, int ldc, Zc* work) {
  if (!znonzero(tau)) return;
  if (left) {
    zgemv_c(m, n, c, ldc, v, incv, work);
    zgerc(m, n, zneg(tau), v, incv, work, 1, c, ldc);
  } else {
    zgemv_n(m, n, c, ldc, v, incv, work);
    zgerc(m, n, zneg(tau), work, 1, v, incv, c, ldc);
  }
}

// Bidiagonal reduction Q^H A P = B, real upper bidiagonal for m >= n (ZGEBD2).
void zgebd2(int m, int n, Zc* a, int lda, double* d, double* e, Zc* tauq, Zc* taup, Zc* work) {
  auto A = [&](int i, int j) -> Zc& { return a[i + (size_t)j * lda]; };
  for (int i = 0; i < n; ++i) {
    Zc alpha = A(i, i);
    zlarfg(m - i, alpha, &A(std::min(i + 1, m - 1), i), 1, tauq[i]);
    d[i] = alpha.re;
    A(i, i) = kOne;
    zlarf(true, m - i, n - i - 1, &A(i, i), 1, zconj(tauq[i]), &A(i, i + 1 < n ? i + 1 : i), lda, work);
    A(i, i) = Zc{d[i], 0.0};
    if (i < n - 1) {
      zlacgv(n - i - 1, &A(i, i + 1), lda);
      alpha = A(i, i + 1);
      zlarfg(n - i - 1, alpha, &A(i, std::min(i + 2, n - 1)), lda, taup[i]);
      e[i] = alpha.re;
      A(i, i + 1) = kOne;
      zlarf(false, m - i - 1, n - i - 1, &A(i, i + 1), lda, taup[i], &A(i + 1 < m ? i + 1 : i, i + 1), lda, work);
      zlacgv(n - i - 1, &A(i, i + 1), lda);
      A(i, i + 1) = Zc{e[i], 0.0};
    } else {
      taup[i] = kZero;
    }
  }
}

// Q with orthonormal columns from k reflectors stored in the columns (ZUNG2R).
void zung2r(int m, int n, int k, Zc* a, int lda, const Zc* tau, Zc* work) {
  if (n <= 0) return;
  auto A = [&](int i, int j) -> Zc& { return a[i + (size_t)j * lda]; };
  for (int j = k; j < n; ++j) {
    for (int l = 0; l < m; ++l) A(l, j) = kZero;
    A(j, j) = kOne;
  }
  for (int i = k - 1; i >= 0; --i) {
    if (i < n - 1) {
      A(i, i) = kOne;
      zlarf(true, m - i, n - i - 1, &A(i, i), 1, tau[i], &A(i, i + 1), lda, work);
    }
    if (i < m - 1) zscal(m - i - 1, zneg(tau[i]), &A(i + 1, i), 1);
    A(i, i) = zsub(kOne, tau[i]);
    for (int l = 0; l < i; ++l) A(l, i) = kZero;
  }
}

// Q with orthonormal rows from k reflectors stored in the rows (ZUNGL2).
void zungl2(int m, int n, int k, Zc* a, int lda, const Zc* tau, Zc* work) {
  if (m <= 0) return;
  auto A = [&](int i, int j) -> Zc& { return a[i + (size_t)j * lda]; };
  if (k < m) {
    for (int j = 0; j < n; ++j) {
      for (int l = k; l < m; ++l) A(l, j) = kZero;
      if (j >= k && j < m) A(j, j) = kOne;
    }
  }
  for (int i = k - 1; i >= 0; --i) {
    if (i < n - 1) {
      zlacgv(n - i - 1, &A(i, i + 1), lda);
      if (i < m - 1) {
        A(i, i) = kOne;
        zlarf(false, m - i - 1, n - i, &A(i, i), lda, zconj(tau[i]), &A(i + 1, i), lda, work);
      }
      zscal(n - i - 1, zneg(tau[i]), &A(i, i + 1), lda);
      zlacgv(n - i - 1, &A(i, i + 1), lda);
    }
    A(i, i) = zsub(kOne, zconj(tau[i]));
    for (int l = 0; l < i; ++l) A(i, l) = kZero;
  }
}

// P^H of the bidiagonal reduction of a square n x n matrix (ZUNGBR 'P' with
// k >= n: shift the reflectors one row down, then ZUNGLQ -> ZUNGL2).
void zungbr_p_square(int n, Zc* a, int lda, const Zc* tau, Zc* work) {
  if (n == 0) return;
  auto A = [&](int i, int j) -> Zc& { return a[i + (size_t)j * lda]; };
  A(0, 0) = kOne;
  for (int i = 1; i < n; ++i) A(i, 0) = kZero;
  for (int j = 1; j < n; ++j) {
    for (int i = j - 1; i >= 1; --i) A(i, j) = A(i - 1, j);
    A(0, j) = kZero;
  }
  if (n > 1) zungl2(n - 1, n - 1, n - 1, &A(1, 1), lda, tau, work);
}

}  // namespace

int dsyev_vu(int n, double* a, int lda, double* w, double* work, int lwork) {
  if (n < 0 || lda < std::max(1, n) || lwork < std::max(1, 3 * n - 1)) return -1;
  if (n == 0) return 0;
  if (n == 1) {
    w[0] = a[0];
    a[0] = 1.0;
    return 0;
  }
  const double safmin = kSafmin, eps = kPrec;
  const double smlnum = safmin / eps, bignum = 1.0 / smlnum;
  const double rmin = std::sqrt(smlnum), rmax = std::sqrt(bignum);
  double anrm = 0.0;  // DLANSY 'M', upper triangle
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= j; ++i) anrm = std::max(anrm, std::fabs(a[i + (size_t)j * lda]));
  int iscale = 0;
  double sigma = 1.0;
  if (anrm > 0.0 && anrm < rmin) {
    iscale = 1;
    sigma = rmin / anrm;
  } else if (anrm > rmax) {
    iscale = 1;
    sigma = rmax / anrm;
  }
  if (iscale == 1) dlascl(2, 1.0, sigma, n, n, a, lda);
  double* e = work;
  double* tau = work + n;
  double* wrk = work + 2 * n;
  dsytd2_u(n, a, lda, w, e, tau);
  dorgtr_u(n, a, lda, tau, wrk);
  const int info = dsteqr_v(n, w, e, a, lda, tau);
  if (iscale == 1) {
    const int imax = info == 0 ? n : info - 1;
    dscal(imax, 1.0 / sigma, w, 1);
  }
  return info;
}

int dgesvd_aa(int n, double* a, int lda, double* s, double* u, int ldu, double* vt, int ldvt, double* work,
              int lwork) {
  if (n < 0 || lda < std::max(1, n) || ldu < std::max(1, n) || ldvt < std::max(1, n)) return -1;
  if (lwork < std::max(3 * n + n, 5 * n)) return -13;
  if (n == 0) return 0;
  if (n == 1) {
    // the QR path (m >= int(1.6 n) for n = 1): Q = 1, R = a, DBDSQR flips the sign into VT
    const double d = a[0];
    u[0] = 1.0;
    vt[0] = 1.0;
    s[0] = d;
    if (d < 0.0) {
      s[0] = -d;
      vt[0] = -1.0;
    }
    return 0;
  }
  const double eps = kPrec;
  const double smlnum = std::sqrt(kSafmin) / eps, bignum = 1.0 / smlnum;
  double anrm = 0.0;  // DLANGE 'M'
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) anrm = std::max(anrm, std::fabs(a[i + (size_t)j * lda]));
  int iscl = 0;
  if (anrm > 0.0 && anrm < smlnum) {
    iscl = 1;
    dlascl(0, anrm, smlnum, n, n, a, lda);
  } else if (anrm > bignum) {
    iscl = 1;
    dlascl(0, anrm, bignum, n, n, a, lda);
  }
  double* e = work;
  double* tauq = work + n;
  double* taup = work + 2 * n;
  double* wrk = work + 3 * n;
  dgebd2(n, n, a, lda, s, e, tauq, taup, wrk);
  for (int j = 0; j < n; ++j)  // DLACPY 'L'
    for (int i = j; i < n; ++i) u[i + (size_t)j * ldu] = a[i + (size_t)j * lda];
  dorgbr_q(n, n, n, u, ldu, tauq, wrk);
  for (int j = 0; j < n; ++j)  // DLACPY 'U'
    for (int i = 0; i <= j; ++i) vt[i + (size_t)j * ldvt] = a[i + (size_t)j * lda];
  dorgbr_p_square(n, vt, ldvt, taup, wrk);
  const int info = bdsqr_u(n, n, n, s, e, vt, ldvt, u, ldu, work + n);
  if (iscl == 1) {
    if (anrm > bignum) dlascl(0, bignum, anrm, n, 1, s, n);
    if (anrm < smlnum) dlascl(0, smlnum, anrm, n, 1, s, n);
  }
  return info;
}

int zgesvd_aa(int n, Zc* a, int lda, double* s, Zc* u, int ldu, Zc* vt, int ldvt, Zc* work, int lwork,
              double* rwork) {
  if (n < 0 || lda < std::max(1, n) || ldu < std::max(1, n) || ldvt < std::max(1, n)) return -1;
  if (lwork < std::max(1, 3 * n)) return -13;
  if (n == 0) return 0;
  // n = 1 takes the QR path (m >= int(1.6 n)), which is not restated: the
  // reference's complex factorizations treat 1 x 1 blocks themselves
  // (ilu2cs.cpp:137-146, 730-739).
  if (n == 1) return -1;
  const double eps = kPrec;
  const double smlnum = std::sqrt(kSafmin) / eps, bignum = 1.0 / smlnum;
  double anrm = 0.0;  // ZLANGE 'M': the largest modulus
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      const Zc v = a[i + (size_t)j * lda];
      anrm = std::max(anrm, std::hypot(v.re, v.im));
    }
  // ZLASCL 'G' scales both parts of every entry by the same real factors:
  // DLASCL over the interleaved doubles
  double* ad = reinterpret_cast<double*>(a);
  int iscl = 0;
  if (anrm > 0.0 && anrm < smlnum) {
    iscl = 1;
    dlascl(0, anrm, smlnum, 2 * n, n, ad, 2 * lda);
  } else if (anrm > bignum) {
    iscl = 1;
    dlascl(0, anrm, bignum, 2 * n, n, ad, 2 * lda);
  }
  double* e = rwork;
  Zc* tauq = work;
  Zc* taup = work + n;
  Zc* wrk = work + 2 * n;
  zgebd2(n, n, a, lda, s, e, tauq, taup, wrk);
  for (int j = 0; j < n; ++j)  // ZLACPY 'L'
    for (int i = j; i < n; ++i) u[i + (size_t)j * ldu] = a[i + (size_t)j * lda];
  zung2r(n, n, n, u, ldu, tauq, wrk);  // ZUNGBR 'Q' with m >= k -> ZUNGQR
  for (int j = 0; j < n; ++j)  // ZLACPY 'U'
    for (int i = 0; i <= j; ++i) vt[i + (size_t)j * ldvt] = a[i + (size_t)j * lda];
  zungbr_p_square(n, vt, ldvt, taup, wrk);
  const int info = bdsqr_u(n, n, n, s, e, vt, ldvt, u, ldu, rwork + n);
  if (iscl == 1) {
    if (anrm > bignum) dlascl(0, bignum, anrm, n, 1, s, n);
    if (anrm < smlnum) dlascl(0, smlnum, anrm, n, 1, s, n);
  }
  return info;
}

}  // namespace lapack30
}  // namespace bcm
