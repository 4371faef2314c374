/*
 * ORACLE - test infrastructure only. Plain-C restatement of the solve phase
 * of Bslib (the reference's serial CPU path) used to check libbcmhip's GPU
 * kernels and drivers. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / the timed CPU
 * baseline; the product never calls it.
 *
 * Matrices: block CSR with block-size prefix sums sprnd[nsup+1], ia, ja,
 * bsa (value offset of each block) and column-major blocks
 * (element (r,c) of block j at a[bsa[j] + c*bla_i + r], mvm.cpp:2560).
 * Vectors: column-major n x nrhs.
 *
 * Parity pins (tests/test_oracle.py): every operator is checked against the
 * compiled reference sources (oracle/_ref, built from Bslib/mvm.cpp,
 * svector.cpp, cg.cpp, lanczos.cpp ...) bit-for-bit, and against dense
 * closed forms; orc_bicgstab and orc_fgmres have no reference counterpart and
 * are pinned by closed forms (exact preconditioner -> 1 iteration, dense
 * solve agreement, monotone FGMRES residuals). orc_block_gmres restates
 * gmres.cpp, whose build needs the Fortran dlarfg (lapackdm.f) and the MPI
 * tree of CQrd, so it is not compiled from the reference; it is pinned by
 * brute force: each step's iterate and residual norms equal the dense
 * least-squares minimiser over the block Krylov space.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int nsup;
  const int* sprnd;
  const int* ia;
  const int* ja;
  const long long* bsa;
  const double* a;
} orc_bsr;

#define BLA(m, i) ((m)->sprnd[(i) + 1] - (m)->sprnd[(i)])

/* y = A x, CSMatrixRS::MvmA (mvm.cpp:2530-2569): per block row, per block,
 * y_i[r] += a(r,c) x_j[c] in the order r outer, c inner. */
void orc_mvma(const orc_bsr* m, const double* x, double* y, int nrhs) {
  const int n = m->sprnd[m->nsup];
  memset(y, 0, sizeof(double) * (size_t)n * nrhs);
  for (int q = 0; q < nrhs; ++q)
    for (int i = 0; i < m->nsup; ++i) {
      const int bi = BLA(m, i), oi = m->sprnd[i];
      for (int j = m->ia[i]; j < m->ia[i + 1]; ++j) {
        const int c0 = m->ja[j], bj = BLA(m, c0), oj = m->sprnd[c0];
        const double* blk = m->a + m->bsa[j];
        for (int r = 0; r < bi; ++r) {
          double* py = y + (size_t)q * n + oi + r;
          for (int c = 0; c < bj; ++c) *py += blk[(size_t)c * bi + r] * x[(size_t)q * n + oj + c];
        }
      }
    }
}

/* y = A^T x, CSMatrixRS::MvmAt (mvm.cpp:2593-2632). */
void orc_mvmat(const orc_bsr* m, const double* x, double* y, int nrhs) {
  const int n = m->sprnd[m->nsup];
  memset(y, 0, sizeof(double) * (size_t)n * nrhs);
  for (int q = 0; q < nrhs; ++q)
    for (int i = 0; i < m->nsup; ++i) {
      const int bi = BLA(m, i), oi = m->sprnd[i];
      for (int j = m->ia[i]; j < m->ia[i + 1]; ++j) {
        const int c0 = m->ja[j], bj = BLA(m, c0), oj = m->sprnd[c0];
        const double* blk = m->a + m->bsa[j];
        for (int r = 0; r < bi; ++r) {
          const double xv = x[(size_t)q * n + oi + r];
          for (int c = 0; c < bj; ++c) y[(size_t)q * n + oj + c] += blk[(size_t)c * bi + r] * xv;
        }
      }
    }
}

/* y = A x from upper storage, CSMatrixRS::MvmASymm (mvm.cpp:3008-3037). */
void orc_mvmasymm(const orc_bsr* m, const double* x, double* y) {
  const int n = m->sprnd[m->nsup];
  memset(y, 0, sizeof(double) * (size_t)n);
  for (int i = 0; i < m->nsup; ++i) {
    const int bi = BLA(m, i), oi = m->sprnd[i];
    for (int j = m->ia[i]; j < m->ia[i + 1]; ++j) {
      const int c0 = m->ja[j], bj = BLA(m, c0), oj = m->sprnd[c0];
      const double* blk = m->a + m->bsa[j];
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bj; ++c) y[oi + r] += blk[(size_t)c * bi + r] * x[oj + c];
      if (i != c0)
        for (int r = 0; r < bj; ++r)
          for (int c = 0; c < bi; ++c) y[oj + r] += blk[(size_t)r * bi + c] * x[oi + c];
    }
  }
}

/* Forward solve with a factor stored by columns, CSMatrixRS::SolveL
 * (mvm.cpp:3140-3217): x_i <- D_i^T x_i (D_i the stored inverse diagonal
 * block), then x_k -= F_ik^T x_i for the blocks of row i. */
void orc_solvel(const orc_bsr* f, const double* x, double* lx, int nrhs) {
  const int n = f->sprnd[f->nsup];
  int bmax = 1;
  for (int i = 0; i < f->nsup; ++i) bmax = BLA(f, i) > bmax ? BLA(f, i) : bmax;
  double* v = (double*)malloc(sizeof(double) * bmax);
  memcpy(lx, x, sizeof(double) * (size_t)n * nrhs);
  for (int q = 0; q < nrhs; ++q) {
    double* y = lx + (size_t)q * n;
    for (int i = 0; i < f->nsup; ++i) {
      const int bi = BLA(f, i), oi = f->sprnd[i];
      const double* d = f->a + f->bsa[f->ia[i]];
      for (int r = 0; r < bi; ++r) {
        double s = 0.0;
        for (int c = 0; c < bi; ++c) s += d[(size_t)r * bi + c] * y[oi + c];
        v[r] = s;
      }
      memcpy(y + oi, v, sizeof(double) * bi);
      for (int j = f->ia[i] + 1; j < f->ia[i + 1]; ++j) {
        const int k = f->ja[j], bk = BLA(f, k), ok = f->sprnd[k];
        const double* blk = f->a + f->bsa[j];
        for (int r = 0; r < bk; ++r)
          for (int c = 0; c < bi; ++c) y[ok + r] -= blk[(size_t)r * bi + c] * y[oi + c];
      }
    }
  }
  free(v);
}

/* Backward solve with a factor stored by rows, CSMatrixRS::SolveU
 * (mvm.cpp:3410-3484): v = x_i - sum_j F_ij y_j, y_i = D_i v. */
void orc_solveu(const orc_bsr* f, const double* x, double* ux, int nrhs) {
  const int n = f->sprnd[f->nsup];
  int bmax = 1;
  for (int i = 0; i < f->nsup; ++i) bmax = BLA(f, i) > bmax ? BLA(f, i) : bmax;
  double* v = (double*)malloc(sizeof(double) * bmax);
  for (int q = 0; q < nrhs; ++q) {
    const double* xx = x + (size_t)q * n;
    double* y = ux + (size_t)q * n;
    for (int i = f->nsup - 1; i >= 0; --i) {
      const int bi = BLA(f, i), oi = f->sprnd[i];
      memcpy(v, xx + oi, sizeof(double) * bi);
      for (int j = f->ia[i] + 1; j < f->ia[i + 1]; ++j) {
        const int k = f->ja[j], bk = BLA(f, k), ok = f->sprnd[k];
        const double* blk = f->a + f->bsa[j];
        for (int r = 0; r < bi; ++r)
          for (int c = 0; c < bk; ++c) v[r] -= blk[(size_t)c * bi + r] * y[ok + c];
      }
      const double* d = f->a + f->bsa[f->ia[i]];
      for (int r = 0; r < bi; ++r) {
        double s = 0.0;
        for (int c = 0; c < bi; ++c) s += d[(size_t)c * bi + r] * v[c];
        y[oi + r] = s;
      }
    }
  }
  free(v);
}

/* ScProdSVect (svector.cpp:364-371): first right-hand side only. */
double orc_scprod(int n, const double* x, const double* y) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s = s + x[i] * y[i];
  return s;
}

/* MaxProdSVect (svector.cpp:376-385). */
double orc_maxprod(int n, const double* x, const double* w) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double a = x[i] * w[i];
    if (a < 0) a = -a;
    if (s < a) s = a;
  }
  return s;
}

typedef struct {
  int nitperf;    /* reference's k at exit */
  int updates;    /* number of solution updates (A applications inside the loop) */
  double resfin;  /* last tested quantity */
  int breakdown;
} orc_result;

/* Preconditioned CG, CSVector::Cg serial form (cg.cpp:31-349), with the
 * factor used as SolveL then SolveU (cg.cpp:141-143) and the stopping rules
 * of cg.cpp:238-241 (ittype 0: ||r||^2 <= eps^2 ||r0||^2, 1: absolute,
 * 2: max|r_i w_i| <= eps); ittype 3 stops on ||r|| <= eps ||b||. Returns the
 * stored good solution (cg.cpp:194-207). The matrix is upper-stored (MvmASymm). */
orc_result orc_cg(const orc_bsr* a, const orc_bsr* u, const double* x0, const double* b, int niter, double eps,
                  int ittype, double* sol_out) {
  const int n = a->sprnd[a->nsup];
  const size_t vb = sizeof(double) * (size_t)n;
  double *sol = malloc(vb), *bloc = malloc(vb), *r = malloc(vb), *p = malloc(vb), *q = malloc(vb), *z = malloc(vb),
         *good = malloc(vb), *w = malloc(vb);
  orc_result res = {0, 0, 0.0, 0};
  for (int i = 0; i < n; ++i) w[i] = 1.0;
  memcpy(sol, x0, vb);
  memcpy(good, x0, vb);
  orc_mvmasymm(a, sol, r);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  memcpy(bloc, r, vb);
  const double bnorm2 = orc_scprod(n, b, b);
  double errmin = ittype == 2 ? orc_maxprod(n, r, w) : 0.0;
  double alpha = 1.0, beta = 0.0, gamold = 0.0, enrsq0 = 0.0;
  memset(p, 0, vb);
  for (int k = 1; k < niter; ++k) {
    orc_solvel(u, r, q, 1);
    orc_solveu(u, q, z, 1);
    const double enrsq = orc_scprod(n, r, r);
    const double gamma = orc_scprod(n, r, z);
    if (k == 1) enrsq0 = enrsq;
    double errloc = ittype == 2 ? orc_maxprod(n, r, w) : 0.0;
    memcpy(good, sol, vb);
    if (ittype == 2) errmin = errloc;
    res.nitperf = k;
    res.resfin = ittype == 2 ? errmin : sqrt(enrsq);
    const int it = k - 1;
    if (it > niter) break;
    if (ittype == 0 && enrsq <= eps * eps * enrsq0) break;
    if (ittype == 1 && enrsq <= eps * eps) break;
    if (ittype == 2 && errmin <= fabs(eps)) break;
    if (ittype == 3 && enrsq <= eps * eps * bnorm2) break;
    if (k > 1) beta = gamma / gamold;
    for (int i = 0; i < n; ++i) p[i] *= beta;
    for (int i = 0; i < n; ++i) p[i] += 1.0 * z[i];
    orc_mvmasymm(a, p, z);
    alpha = orc_scprod(n, p, z);
    alpha = gamma / alpha;
    if (alpha < 0.0) {
      res.breakdown = 1;
      break;
    }
    gamold = gamma;
    for (int i = 0; i < n; ++i) r[i] += -alpha * z[i];
    for (int i = 0; i < n; ++i) sol[i] += alpha * p[i];
    res.updates++;
  }
  memcpy(sol_out, good, vb);
  free(sol); free(bloc); free(r); free(p); free(q); free(z); free(good); free(w);
  return res;
}

/* Parallel-form CG, CSVector::Cg(tree, CMVM, CSOLVELU) (cg.cpp:355-771) on
 * one process: r = A x - b, u = M^-1 r; per iteration uAu = (Au, u), ru =
 * (r, u), acoef = -ru / uAu, r += acoef Au, the stopping test on ||r||
 * (sttype 0: ||r|| <= eps, 1: ||r|| <= eps ||r0||), then v = M^-1 r,
 * bcoef = -(v, Au) / uAu, x += acoef u, u = v + bcoef u. The reference leaves
 * the loop at the stopping test before the solution takes the last acoef u
 * step (cg.cpp:607-614 against :660-661); apply_last = 1 applies that step at
 * exit. nitperf = k - 1 at exit; A is upper-stored (MvmASymm), M = U^T U. */
orc_result orc_cg_par(const orc_bsr* a, const orc_bsr* u, const double* x0, const double* b, int niter, double eps,
                      int sttype, int apply_last, double* sol_out) {
  const int n = a->sprnd[a->nsup];
  const size_t vb = sizeof(double) * (size_t)n;
  double *sol = malloc(vb), *r = malloc(vb), *uu = malloc(vb), *au = malloc(vb), *v = malloc(vb), *t = malloc(vb);
  orc_result res = {0, 0, 0.0, 0};
  memcpy(sol, x0, vb);
  orc_mvmasymm(a, sol, r);
  for (int i = 0; i < n; ++i) r[i] = r[i] - b[i];
  double aux = 0.0;
  for (int i = 0; i < n; ++i) aux += r[i] * r[i];
  const double resi0 = sqrt(aux);
  orc_solvel(u, r, t, 1);
  orc_solveu(u, t, uu, 1);
  for (int k = 1; k < niter; ++k) {
    const int it = k - 1;
    orc_mvmasymm(a, uu, au);
    const double uau = orc_scprod(n, au, uu);
    const double ru = orc_scprod(n, r, uu);
    if (uau < 0.0) { res.breakdown = 1; break; }
    const double acoef = (ru * -1.0) / uau;
    for (int i = 0; i < n; ++i) r[i] = r[i] + acoef * au[i];
    aux = 0.0;
    for (int i = 0; i < n; ++i) aux += r[i] * r[i];
    const double resi = sqrt(aux);
    res.updates = k;
    if ((sttype == 0 && resi <= eps) || (sttype == 1 && resi <= eps * resi0)) {
      res.nitperf = it;
      if (apply_last)
        for (int i = 0; i < n; ++i) sol[i] = sol[i] + acoef * uu[i];
      break;
    }
    orc_solvel(u, r, t, 1);
    orc_solveu(u, t, v, 1);
    const double vau = orc_scprod(n, v, au);
    const double bcoef = (vau * -1.0) / uau;
    for (int i = 0; i < n; ++i) {
      sol[i] = sol[i] + acoef * uu[i];
      uu[i] = v[i] + bcoef * uu[i];
    }
    res.nitperf = it;
  }
  orc_mvmasymm(a, sol, r);
  aux = 0.0;
  for (int i = 0; i < n; ++i) aux += (r[i] - b[i]) * (r[i] - b[i]);
  res.resfin = sqrt(aux);
  memcpy(sol_out, sol, vb);
  free(sol); free(r); free(uu); free(au); free(v); free(t);
  return res;
}

/* Two-sided Lanczos (BiCG-type), CSVector::Lanczos serial form
 * (lanczos.cpp:34-371): preconditioner L.SolveL then U.SolveU, transposed
 * preconditioner U.SolveL then L.SolveU, A and A^T products; stops on
 * sttype (0 absolute 2-norm, 1 relative to ||r0||, 2 max-norm) or breakdown
 * (|delDel| < 1e-100 twice); returns the best iterate by max-norm. */
orc_result orc_lanczos(const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, const double* x0, const double* b,
                       int niter, double eps, int sttype, double* sol_out) {
  const int n = a->sprnd[a->nsup];
  const size_t vb = sizeof(double) * (size_t)n;
  double *sol = malloc(vb), *r = malloc(vb), *good = malloc(vb), *sol1 = calloc(n, 8), *Delta = malloc(vb),
         *Delta1 = calloc(n, 8), *delta = malloc(vb), *delta1 = calloc(n, 8), *Img = malloc(vb), *p = malloc(vb),
         *w = malloc(vb);
  orc_result res = {0, 0, 0.0, 0};
  for (int i = 0; i < n; ++i) w[i] = 1.0;
  memcpy(sol, x0, vb);
  orc_mvma(a, sol, r, 1);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  double errmin = orc_maxprod(n, r, w);
  const double resini = sqrt(orc_scprod(n, r, r));
  memcpy(good, sol, vb);
  memcpy(Delta, r, vb);
  orc_solvel(l, r, p, 1);
  orc_solveu(u, p, delta, 1);
  double gamma = 0.0, gamma1, ro = 1.0, delDel = 0.0, delDel1;
  for (int k = 1; k < niter; ++k) {
    orc_mvma(a, delta, r, 1);
    orc_solvel(l, r, p, 1);
    orc_solveu(u, p, Img, 1);
    gamma1 = gamma;
    delDel1 = delDel;
    delDel = orc_scprod(n, delta, Delta);
    const double dtemp = orc_scprod(n, Img, Delta);
    gamma = delDel / dtemp;
    if (k > 1) ro = 1.0 / (1.0 - (gamma / gamma1) * (delDel / delDel1) / ro);
    const double rogamma = ro * gamma, unitro = 1. - ro;
    for (int i = 0; i < n; ++i) p[i] = sol[i] * ro + sol1[i] * unitro;
    memcpy(sol1, sol, vb);
    for (int i = 0; i < n; ++i) sol[i] = p[i] + delta[i] * rogamma;
    orc_mvma(a, sol, r, 1);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const double errloc = orc_maxprod(n, r, w);
    const double resi = sqrt(orc_scprod(n, r, r));
    if (errloc < errmin) {
      memcpy(good, sol, vb);
      errmin = errloc;
    }
    res.nitperf = k;
    res.updates = k;
    res.resfin = sttype == 2 ? errmin : resi;
    const int brk = fabs(delDel) < 1.e-100 && fabs(delDel1) < 1.e-100;
    if (sttype == 2 && (k > niter || errmin <= eps || brk)) { res.breakdown = brk; break; }
    if (sttype == 0 && (k > niter || resi <= eps || brk)) { res.breakdown = brk; break; }
    if (sttype == 1 && (k > niter || resi <= eps * resini || brk)) { res.breakdown = brk; break; }
    for (int i = 0; i < n; ++i) p[i] = delta[i] * ro + delta1[i] * unitro;
    memcpy(delta1, delta, vb);
    for (int i = 0; i < n; ++i) delta[i] = p[i] - Img[i] * rogamma;
    orc_solvel(u, Delta, p, 1);
    orc_solveu(l, p, r, 1);
    orc_mvmat(a, r, Img, 1);
    for (int i = 0; i < n; ++i) p[i] = Delta[i] * ro + Delta1[i] * unitro;
    memcpy(Delta1, Delta, vb);
    for (int i = 0; i < n; ++i) Delta[i] = p[i] - Img[i] * rogamma;
  }
  memcpy(sol_out, good, vb);
  free(sol); free(r); free(good); free(sol1); free(Delta); free(Delta1); free(delta); free(delta1); free(Img);
  free(p); free(w);
  return res;
}

/* Preconditioner z = U^-1 L^-1 r (L.SolveL then U.SolveU). */
static void precond(const orc_bsr* l, const orc_bsr* u, const double* r, double* z, double* tmp) {
  orc_solvel(l, r, tmp, 1);
  orc_solveu(u, tmp, z, 1);
}

/* Right-preconditioned BiCGStab (van der Vorst 1992), the algorithm of
 * libbcmhip's GPU driver, stopping on ||r|| <= eps ||b|| (tested on s and on r).
 * A is full storage; l == u for an IC2 factor. */
orc_result orc_bicgstab(const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, const double* x0, const double* b,
                        int niter, double eps, double* sol_out) {
  const int n = a->sprnd[a->nsup];
  const size_t vb = sizeof(double) * (size_t)n;
  double *x = malloc(vb), *r = malloc(vb), *rh = malloc(vb), *p = calloc(n, 8), *v = calloc(n, 8), *ph = malloc(vb),
         *s = malloc(vb), *sh = malloc(vb), *t = malloc(vb), *tmp = malloc(vb);
  orc_result res = {0, 0, 0.0, 0};
  memcpy(x, x0, vb);
  orc_mvma(a, x, r, 1);
  for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
  memcpy(rh, r, vb);
  const double thr = eps * eps * orc_scprod(n, b, b);
  double rho = orc_scprod(n, rh, r), rho_old = 1.0, alpha = 1.0, omega = 1.0;
  double rr = orc_scprod(n, r, r);
  if (rr <= thr) goto done;
  for (int k = 1; k <= niter; ++k) {
    const double beta = (rho / rho_old) * (alpha / omega);
    for (int i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    precond(l, u, p, ph, tmp);
    orc_mvma(a, ph, v, 1);
    const double rv = orc_scprod(n, rh, v);
    if (rv == 0.0) { res.breakdown = 1; break; }
    alpha = rho / rv;
    for (int i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    const double ss = orc_scprod(n, s, s);
    if (ss <= thr) {
      for (int i = 0; i < n; ++i) x[i] += alpha * ph[i];
      res.nitperf = k;
      res.resfin = sqrt(ss);
      break;
    }
    precond(l, u, s, sh, tmp);
    orc_mvma(a, sh, t, 1);
    const double ts = orc_scprod(n, t, s), tt = orc_scprod(n, t, t);
    if (tt == 0.0) { res.breakdown = 1; break; }
    omega = ts / tt;
    for (int i = 0; i < n; ++i) x[i] = x[i] + alpha * ph[i] + omega * sh[i];
    for (int i = 0; i < n; ++i) r[i] = s[i] - omega * t[i];
    rr = orc_scprod(n, r, r);
    rho_old = rho;
    rho = orc_scprod(n, rh, r);
    res.nitperf = k;
    res.resfin = sqrt(rr);
    if (rr <= thr) break;
    if (rho == 0.0) { res.breakdown = 1; break; }
  }
done:
  res.updates = res.nitperf;
  memcpy(sol_out, x, vb);
  free(x); free(r); free(rh); free(p); free(v); free(ph); free(s); free(sh); free(t); free(tmp);
  return res;
}

/* Flexible GMRES(m) (Saad 1993) with modified Gram-Schmidt and Givens
 * rotations, right preconditioning through the factor (symm: A is
 * upper-stored symmetric and applied as MvmASymm), stopping on
 * ||b - A x|| <= eps ||b|| (the Arnoldi residual estimate, confirmed by the
 * true residual at every restart). */
orc_result orc_fgmres(const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, const double* x0, const double* b,
                      int m, int niter, double eps, int symm, double* sol_out) {
  const int n = a->sprnd[a->nsup];
  const size_t vb = sizeof(double) * (size_t)n;
  double* V = malloc(vb * (m + 1));
  double* Z = malloc(vb * m);
  double* H = calloc((size_t)(m + 1) * m, 8);
  double *cs = malloc(8 * m), *sn = malloc(8 * m), *g = malloc(8 * (m + 1)), *y = malloc(8 * m);
  double *x = malloc(vb), *r = malloc(vb), *tmp = malloc(vb);
  orc_result res = {0, 0, 0.0, 0};
  memcpy(x, x0, vb);
  const double bnorm = sqrt(orc_scprod(n, b, b));
  int total = 0;
  while (total < niter) {
    if (symm) orc_mvmasymm(a, x, r); else orc_mvma(a, x, r, 1);
    for (int i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const double beta = sqrt(orc_scprod(n, r, r));
    res.resfin = beta / bnorm;
    if (beta <= eps * bnorm) break;
    for (int i = 0; i < n; ++i) V[i] = r[i] / beta;
    memset(g, 0, 8 * (m + 1));
    g[0] = beta;
    int j;
    for (j = 0; j < m && total < niter; ++j) {
      precond(l, u, V + (size_t)j * n, Z + (size_t)j * n, tmp);
      double* w = V + (size_t)(j + 1) * n;
      if (symm) orc_mvmasymm(a, Z + (size_t)j * n, w); else orc_mvma(a, Z + (size_t)j * n, w, 1);
      for (int i = 0; i <= j; ++i) {
        const double h = orc_scprod(n, w, V + (size_t)i * n);
        H[(size_t)j * (m + 1) + i] = h;
        for (int k = 0; k < n; ++k) w[k] -= h * V[(size_t)i * n + k];
      }
      const double hn = sqrt(orc_scprod(n, w, w));
      H[(size_t)j * (m + 1) + j + 1] = hn;
      if (hn > 0)
        for (int k = 0; k < n; ++k) w[k] /= hn;
      for (int i = 0; i < j; ++i) {
        const double h0 = H[(size_t)j * (m + 1) + i], h1 = H[(size_t)j * (m + 1) + i + 1];
        H[(size_t)j * (m + 1) + i] = cs[i] * h0 + sn[i] * h1;
        H[(size_t)j * (m + 1) + i + 1] = -sn[i] * h0 + cs[i] * h1;
      }
      const double h0 = H[(size_t)j * (m + 1) + j], h1 = H[(size_t)j * (m + 1) + j + 1];
      const double rr = hypot(h0, h1);
      cs[j] = h0 / rr;
      sn[j] = h1 / rr;
      H[(size_t)j * (m + 1) + j] = rr;
      H[(size_t)j * (m + 1) + j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++total;
      if (fabs(g[j + 1]) <= eps * bnorm) { ++j; break; }
    }
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < j; ++k) s -= H[(size_t)k * (m + 1) + i] * y[k];
      y[i] = s / H[(size_t)i * (m + 1) + i];
    }
    for (int i = 0; i < j; ++i)
      for (int k = 0; k < n; ++k) x[k] += y[i] * Z[(size_t)i * n + k];
  }
  res.nitperf = total;
  res.updates = total;
  memcpy(sol_out, x, vb);
  free(V); free(Z); free(H); free(cs); free(sn); free(g); free(y); free(x); free(r); free(tmp);
  return res;
}

/* ---- block GMRES, CSVector::BlockGMRES (gmres.cpp:337-1038) ----
 *
 * Right-preconditioned block GMRES without restart over nrhs right-hand
 * sides: the block Arnoldi basis is the Q factor of a Householder QR that is
 * extended by nrhs columns per step (UpdateQrd, globals.cpp:531-590), the
 * Krylov directions are Q applied to unit blocks (MvmQBlk,
 * globals.cpp:778-840), the block Hessenberg columns are reduced by block
 * Givens rotations, each the explicit Q of a 2nrhs x 2nrhs Householder QR
 * (BlockGivens / ApplyBlockGivens, gmres.cpp:30-90), and the solution update
 * is a back substitution (SolveTriangular, gmres.cpp:130-160). The
 * reflectors follow LAPACK's dlarfg convention (beta = -sign(alpha) ||x||,
 * v(1) = 1); the safe-minimum rescaling of dlarfg is left out. */

/* dlarfg: H [alpha; x] = [beta; 0] with H = I - tau [1; v][1; v]^T; x is
 * overwritten by v, alpha by beta (k = length including alpha). */
static void orc_larfg(int k, double* alpha, double* x, int incx, double* tau) {
  *tau = 0.0;
  if (k <= 1) return;
  double xn2 = 0.0;
  for (int i = 0; i < k - 1; ++i) xn2 += x[(size_t)i * incx] * x[(size_t)i * incx];
  if (xn2 == 0.0) return;
  const double beta = -copysign(sqrt(*alpha * *alpha + xn2), *alpha);
  *tau = (beta - *alpha) / beta;
  const double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < k - 1; ++i) x[(size_t)i * incx] *= sc;
  *alpha = beta;
}

/* Householder QR of columns [ibeg, iend) of the m-row panel a (ld lda),
 * columns < ibeg already reduced (UpdateQrd, globals.cpp:531-590). */
static void orc_update_qrd(int m, int ibeg, int iend, double* a, int lda, double* tau) {
  for (int i = ibeg; i < iend; ++i) {
    double* ci = a + (size_t)i * lda;
    for (int j = 0; j < i && j < m; ++j) {
      const double* vj = a + (size_t)j * lda;
      double s = ci[j];
      for (int k = j + 1; k < m; ++k) s += ci[k] * vj[k];
      s *= tau[j];
      ci[j] -= s;
      for (int k = j + 1; k < m; ++k) ci[k] -= s * vj[k];
    }
    if (i < m) orc_larfg(m - i, ci + i, ci + i + 1, 1, tau + i);
    else tau[i] = 0.0;
  }
}

/* qx = Q [x; 0] for the first nq reflectors of q (MvmQBlk, globals.cpp:778-840):
 * x is nq x nrhs (ld ldx), qx is m x nrhs (ld ldqx). */
static void orc_mvmq(int m, int nq, int nrhs, const double* q, int ldq, const double* tau, const double* x, int ldx,
                     double* qx, int ldqx) {
  for (int c = 0; c < nrhs; ++c) {
    double* y = qx + (size_t)c * ldqx;
    for (int i = 0; i < m; ++i) y[i] = i < nq ? x[(size_t)c * ldx + i] : 0.0;
    for (int j = nq - 1; j >= 0; --j) {
      const double* v = q + (size_t)j * ldq;
      double s = y[j];
      for (int k = j + 1; k < m; ++k) s += v[k] * y[k];
      s = -s * tau[j];
      y[j] += s;
      for (int k = j + 1; k < m; ++k) y[k] += s * v[k];
    }
  }
}

/* rows [row0, row0 + 2 nb) of the nrhs columns of e (ld lde) <- G^T (those rows),
 * G the explicit 2nb x 2nb rotation (ApplyBlockGivens, gmres.cpp:62-90). */
static void orc_apply_giv(int nb, int row0, const double* G, int nrhs, double* e, int lde) {
  const int nl = 2 * nb;
  double tmp[64];
  for (int c = 0; c < nrhs; ++c) {
    double* col = e + (size_t)c * lde + row0;
    for (int r = 0; r < nl; ++r) {
      double s = 0.0;
      for (int q = 0; q < nl; ++q) s += col[q] * G[(size_t)r * nl + q];
      tmp[r] = s;
    }
    for (int r = 0; r < nl; ++r) col[r] = tmp[r];
  }
}

static void orc_apply_a(const orc_bsr* a, int symm, int nrhs, const double* x, double* y) {
  const int n = a->sprnd[a->nsup];
  if (!symm) { orc_mvma(a, x, y, nrhs); return; }
  for (int c = 0; c < nrhs; ++c) orc_mvmasymm(a, x + (size_t)c * n, y + (size_t)c * n);
}

/* sttype 0: resi_c < eps; otherwise resi_c / resi0_c < eps (gmres.cpp:742-750).
 * x0, b, sol: n x nrhs column-major. hist (niter x nrhs, may be NULL): the
 * recursion's residual norms after each step; fin (nrhs): ||b_c - A x_c||.
 * Returns nitperf = number of block steps, updates = 1 if every column met
 * the test. nrhs <= 32. */
orc_result orc_block_gmres(const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, int symm, int nrhs, const double* x0,
                           const double* b, int niter, double eps, int sttype, double* sol_out, double* hist,
                           double* fin) {
  const int n = a->sprnd[a->nsup];
  const int nb = nrhs, nb2 = nrhs * nrhs;
  int nl = n;
  if ((niter + 1) * nb > nl) nl = (niter + 1) * nb;
  const int mmax = (niter + 1) * nb, ldb = (niter + 1) * nb, ldr = niter * nb;
  double* p = calloc((size_t)nl * mmax, 8);
  double* taup = calloc((size_t)mmax, 8);
  double* giv = calloc((size_t)niter * nb2 * 4 + 1, 8);
  double* bloc = calloc((size_t)ldb * nb, 8);
  double* eloc = calloc((size_t)(mmax + nb) * nb, 8);
  double* rm = calloc((size_t)ldr * ldr + 1, 8);
  double *r = malloc(8 * (size_t)n * nb), *y = malloc(8 * (size_t)n * nb), *w = malloc(8 * (size_t)n * nb);
  double *tmp = malloc(8 * (size_t)n * nb), *yp = malloc(8 * (size_t)nl * nb);
  double resi0[32], resi[32];
  orc_result res = {0, 0, 0.0, 0};
  memcpy(sol_out, x0, 8 * (size_t)n * nb);
  /* initial block residual and its QR */
  orc_apply_a(a, symm, nb, sol_out, r);
  for (size_t i = 0; i < (size_t)n * nb; ++i) r[i] = b[i] - r[i];
  for (int c = 0; c < nb; ++c) memcpy(p + (size_t)c * nl, r + (size_t)c * n, 8 * (size_t)n);
  orc_update_qrd(nl, 0, nb, p, nl, taup);
  for (int c = 0; c < nb; ++c) {
    double s = 0.0;
    for (int i = 0; i <= c; ++i) {
      bloc[(size_t)c * ldb + i] = p[(size_t)c * nl + i];
      s += p[(size_t)c * nl + i] * p[(size_t)c * nl + i];
    }
    resi0[c] = resi[c] = sqrt(s);
  }
  int iter = 0, conv = 0;
  for (int k = 0; k < niter; ++k) {
    iter = k;
    const int ncolp = (k + 1) * nb;
    /* directions: Q e_(k nb + c), then M^-1 and A */
    memset(eloc, 0, 8 * (size_t)ncolp * nb);
    for (int c = 0; c < nb; ++c) eloc[(size_t)c * ncolp + k * nb + c] = 1.0;
    orc_mvmq(nl, ncolp, nb, p, nl, taup, eloc, ncolp, yp, nl);
    for (int c = 0; c < nb; ++c) memcpy(y + (size_t)c * n, yp + (size_t)c * nl, 8 * (size_t)n);
    orc_solvel(l, y, tmp, nb);
    orc_solveu(u, tmp, w, nb);
    memcpy(y, w, 8 * (size_t)n * nb);
    orc_apply_a(a, symm, nb, y, w);
    /* extend the QR by the new block */
    for (int c = 0; c < nb; ++c) {
      double* pc = p + (size_t)(ncolp + c) * nl;
      memcpy(pc, w + (size_t)c * n, 8 * (size_t)n);
      memset(pc + n, 0, 8 * (size_t)(nl - n));
    }
    orc_update_qrd(nl, ncolp, ncolp + nb, p, nl, taup);
    /* block column of the Hessenberg matrix: R part of the new columns */
    const int iend = ncolp + nb;
    memset(eloc, 0, 8 * (size_t)iend * nb);
    for (int c = 0; c < nb; ++c)
      for (int i = 0; i <= ncolp + c; ++i) eloc[(size_t)c * iend + i] = p[(size_t)(ncolp + c) * nl + i];
    for (int g = 0; g < k; ++g) orc_apply_giv(nb, g * nb, giv + (size_t)g * nb2 * 4, nb, eloc, iend);
    /* new block rotation: QR of the 2nb x nb block, explicit Q */
    {
      double blk[64 * 64], tb[64], ident[64 * 64];
      const int n2 = 2 * nb;
      memset(blk, 0, sizeof(double) * n2 * n2);
      for (int c = 0; c < nb; ++c)
        for (int i = 0; i < n2; ++i) blk[c * n2 + i] = eloc[(size_t)c * iend + k * nb + i];
      orc_update_qrd(n2, 0, n2, blk, n2, tb);
      memset(ident, 0, sizeof(double) * n2 * n2);
      for (int i = 0; i < n2; ++i) ident[i * n2 + i] = 1.0;
      orc_mvmq(n2, n2, n2, blk, n2, tb, ident, n2, giv + (size_t)k * nb2 * 4, n2);
    }
    orc_apply_giv(nb, k * nb, giv + (size_t)k * nb2 * 4, nb, eloc, iend);
    for (int c = 0; c < nb; ++c)
      for (int i = 0; i < ncolp; ++i) rm[(size_t)(k * nb + c) * ldr + i] = eloc[(size_t)c * iend + i];
    orc_apply_giv(nb, k * nb, giv + (size_t)k * nb2 * 4, nb, bloc, ldb);
    conv = 0;
    for (int c = 0; c < nb; ++c) {
      double s = 0.0;
      for (int i = 0; i < nb; ++i) s += bloc[(size_t)c * ldb + (k + 1) * nb + i] * bloc[(size_t)c * ldb + (k + 1) * nb + i];
      resi[c] = sqrt(s);
      if (hist) hist[(size_t)k * nb + c] = resi[c];
      if (sttype == 0 ? resi[c] < eps : resi[c] / resi0[c] < eps) ++conv;
    }
    if (conv == nb) break;
  }
  /* back substitution and the solution update */
  const int ncolp = (iter + 1) * nb, lde = ncolp + nb;
  memset(eloc, 0, 8 * (size_t)lde * nb);
  for (int c = 0; c < nb; ++c) {
    double* z = eloc + (size_t)c * lde;
    for (int i = 0; i < ncolp; ++i) z[i] = bloc[(size_t)c * ldb + i];
    for (int j = ncolp - 1; j >= 0; --j) {
      const double zj = z[j] / rm[(size_t)j * ldr + j];
      z[j] = zj;
      for (int i = 0; i < j; ++i) z[i] -= rm[(size_t)j * ldr + i] * zj;
    }
  }
  orc_mvmq(nl, lde, nb, p, nl, taup, eloc, lde, yp, nl);
  for (int c = 0; c < nb; ++c) memcpy(y + (size_t)c * n, yp + (size_t)c * nl, 8 * (size_t)n);
  orc_solvel(l, y, tmp, nb);
  orc_solveu(u, tmp, w, nb);
  for (size_t i = 0; i < (size_t)n * nb; ++i) sol_out[i] += w[i];
  orc_apply_a(a, symm, nb, sol_out, r);
  double rmax = 0.0;
  for (int c = 0; c < nb; ++c) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      const double d = b[(size_t)c * n + i] - r[(size_t)c * n + i];
      s += d * d;
    }
    if (fin) fin[c] = sqrt(s);
    if (sqrt(s) > rmax) rmax = sqrt(s);
  }
  res.nitperf = iter + 1;
  res.updates = conv == nb;
  res.resfin = rmax;
  free(p); free(taup); free(giv); free(bloc); free(eloc); free(rm); free(r); free(y); free(w); free(tmp); free(yp);
  return res;
}

/* ---- block Lanczos (block Golub-Kahan bidiagonalization with block Givens
 *      rotations), CSVector::BlockLanczosRight / Left / Central
 *      (lanczos.cpp:932-1718, 1719-2466, 2467-3225) ----
 *
 * The three variants run the same recurrence on a preconditioned operator
 * Ahat and its transpose: variant 0 (Right) Ahat = A U^-1 L^-1, 1 (Left)
 * Ahat = U^-1 L^-1 A, 2 (Central) Ahat = L^-1 A U^-1, where L is the factor
 * stored by columns and U by rows (SolveL / SolveU as in mvm.cpp; their
 * transposes are U.SolveL and L.SolveU). Blocks are QR-factored by
 * Householder reflectors (QrdBlk, globals.cpp:413-470; the Q block formed by
 * MvmQBlk, globals.cpp:717-776; R by GetRPartQrd, globals.cpp:920-933); the
 * block Givens rotation is the explicit Q of the 2nrhs x 2nrhs QR of
 * [rhobar; beta] (lanczos.cpp:1244-1268). The reference inverts rho through
 * an SVD (dgesvd, lanczos.cpp:1430-1462); this restatement inverts it by
 * Gauss-Jordan elimination with partial pivoting (equal up to rounding).
 * Convergence is tested every ichk iterations on the true residual
 * b - A x_k of every column (relative to its initial norm), and for Right
 * also on ||Ahat^T r_k|| (lanczos.cpp:1536-1585). Returns nitperf = the number
 * of iterations done, updates = 1 if every column met the test, resfin = the
 * largest ||b_c - A x_c|| / ||b_c - A x0_c||. nrhs <= 16. */
static int orc_inv_small(int nb, const double* a, double* inv) {
  double m[16 * 32];
  for (int r = 0; r < nb; ++r)
    for (int c = 0; c < 2 * nb; ++c) m[r * 2 * nb + c] = c < nb ? a[(size_t)c * nb + r] : (c - nb == r ? 1.0 : 0.0);
  for (int c = 0; c < nb; ++c) {
    int piv = c;
    for (int r = c + 1; r < nb; ++r)
      if (fabs(m[r * 2 * nb + c]) > fabs(m[piv * 2 * nb + c])) piv = r;
    if (m[piv * 2 * nb + c] == 0.0) return 1;
    if (piv != c)
      for (int k = 0; k < 2 * nb; ++k) {
        const double t = m[c * 2 * nb + k];
        m[c * 2 * nb + k] = m[piv * 2 * nb + k];
        m[piv * 2 * nb + k] = t;
      }
    const double d = m[c * 2 * nb + c];
    for (int k = 0; k < 2 * nb; ++k) m[c * 2 * nb + k] /= d;
    for (int r = 0; r < nb; ++r) {
      if (r == c) continue;
      const double f = m[r * 2 * nb + c];
      if (f != 0.0)
        for (int k = 0; k < 2 * nb; ++k) m[r * 2 * nb + k] -= f * m[c * 2 * nb + k];
    }
  }
  for (int r = 0; r < nb; ++r)
    for (int c = 0; c < nb; ++c) inv[(size_t)c * nb + r] = m[r * 2 * nb + nb + c];
  return 0;
}

/* y = Ahat x (transpose = 0) or Ahat^T x, nb columns; t1, t2 scratch */
static void orc_blz_op(int variant, int transpose, const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, int nb,
                       const double* x, double* y, double* t1, double* t2) {
  if (!transpose) {
    if (variant == 0) { orc_solvel(l, x, t1, nb); orc_solveu(u, t1, t2, nb); orc_mvma(a, t2, y, nb); }
    else if (variant == 1) { orc_mvma(a, x, t1, nb); orc_solvel(l, t1, t2, nb); orc_solveu(u, t2, y, nb); }
    else { orc_solveu(u, x, t1, nb); orc_mvma(a, t1, t2, nb); orc_solvel(l, t2, y, nb); }
  } else {
    if (variant == 0) { orc_mvmat(a, x, t1, nb); orc_solvel(u, t1, t2, nb); orc_solveu(l, t2, y, nb); }
    else if (variant == 1) { orc_solvel(u, x, t1, nb); orc_solveu(l, t1, t2, nb); orc_mvmat(a, t2, y, nb); }
    else { orc_solveu(l, x, t1, nb); orc_mvmat(a, t1, t2, nb); orc_solvel(u, t2, y, nb); }
  }
}

/* y += s * X * c (DaxpySVectBlk, svector.cpp:328-343): X is n x nb, c nb x nb column-major */
static void orc_daxpy_blk(int n, int nb, double s, const double* c, const double* x, double* y) {
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < n; ++j) {
      double aux = 0.0;
      for (int k = 0; k < nb; ++k) aux += x[(size_t)k * n + j] * c[i * nb + k];
      y[(size_t)i * n + j] += s * aux;
    }
}

/* Householder QR of the n x nb block w in place; q = its Q block, r = R (nb x nb). */
static void orc_blz_qr(int n, int nb, double* w, double* tau, double* q, double* r) {
  double ident[16 * 16];
  orc_update_qrd(n, 0, nb, w, n, tau);
  memset(ident, 0, sizeof(ident));
  for (int i = 0; i < nb; ++i) ident[i * nb + i] = 1.0;
  orc_mvmq(n, nb, nb, w, n, tau, ident, nb, q, n);
  for (int c = 0; c < nb; ++c)
    for (int i = 0; i < nb; ++i) r[c * nb + i] = i <= c ? w[(size_t)c * n + i] : 0.0;
}

static void orc_tr_small(int nb, double* a) {
  double t[256];
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) t[j * nb + i] = a[i * nb + j];
  memcpy(a, t, sizeof(double) * nb * nb);
}

/* out = a b for nb x nb column-major matrices (lanczos.cpp:1380-1421) */
static void orc_ab(int nb, const double* a, const double* b, double* out) {
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < nb; ++j) {
      double aux = 0.0;
      for (int k = 0; k < nb; ++k) aux += a[k * nb + i] * b[j * nb + k];
      out[j * nb + i] = aux;
    }
}

orc_result orc_block_lanczos(int variant, const orc_bsr* a, const orc_bsr* l, const orc_bsr* u, int nrhs,
                             const double* x0, const double* b, int niter, double eps, int ichk, double* sol_out) {
  const int n = a->sprnd[a->nsup], nb = nrhs, nb2 = nb * nb;
  const size_t vb = sizeof(double) * (size_t)n * nb;
  double *r = malloc(vb), *uu = malloc(vb), *v = malloc(vb), *w = malloc(vb), *qw = malloc(vb), *pw = malloc(vb);
  double *t1 = malloc(vb), *t2 = malloc(vb), *soli = calloc((size_t)n * nb, 8), *xk = malloc(vb);
  double alpha[256], beta[256], rhobar[256], phibar[256], rho[256], rinv[256], theta[256], phi[256], cc[256], ss[256],
      ccbar[256], ssbar[256], coef[256], tau[32], giv[1024], gq[1024], gr[1024], ident[1024];
  double resi0[16], resort0[16];
  orc_result res = {0, 0, 0.0, 0};
  memcpy(sol_out, x0, vb);
  orc_mvma(a, x0, r, nb);
  for (size_t i = 0; i < (size_t)n * nb; ++i) r[i] = b[i] - r[i];
  if (variant != 0) {
    for (int c = 0; c < nb; ++c) {
      double s = 0.0;
      for (int i = 0; i < n; ++i) s += r[(size_t)c * n + i] * r[(size_t)c * n + i];
      resi0[c] = sqrt(s);
    }
    if (variant == 1) { orc_solvel(l, r, t1, nb); orc_solveu(u, t1, r, nb); }
    else { orc_solvel(l, r, t1, nb); memcpy(r, t1, vb); }
  }
  orc_blz_qr(n, nb, r, tau, uu, beta);
  if (variant == 0)
    for (int c = 0; c < nb; ++c) {
      double s = 0.0;
      for (int i = 0; i < nb; ++i) s += beta[c * nb + i] * beta[c * nb + i];
      resi0[c] = sqrt(s);
    }
  orc_blz_op(variant, 1, a, l, u, nb, uu, w, t1, t2);
  orc_blz_qr(n, nb, w, tau, v, alpha);
  orc_tr_small(nb, alpha);
  if (variant == 0)
    for (int c = 0; c < nb; ++c) {
      double s = 0.0;
      for (int i = 0; i < nb; ++i) s += alpha[c * nb + i] * alpha[c * nb + i];
      resort0[c] = sqrt(s);
    }
  memcpy(w, v, vb);
  memcpy(phibar, beta, sizeof(double) * nb2);
  memcpy(rhobar, alpha, sizeof(double) * nb2);
  int k = 0, conv = 0;
  for (k = 0; k < niter; ++k) {
    orc_blz_op(variant, 0, a, l, u, nb, v, qw, t1, t2);
    orc_daxpy_blk(n, nb, -1.0, alpha, uu, qw);
    orc_blz_qr(n, nb, qw, tau, uu, beta);
    orc_blz_op(variant, 1, a, l, u, nb, uu, pw, t1, t2);
    memcpy(coef, beta, sizeof(double) * nb2);
    orc_tr_small(nb, coef);
    orc_daxpy_blk(n, nb, -1.0, coef, v, pw);
    orc_blz_qr(n, nb, pw, tau, v, alpha);
    orc_tr_small(nb, alpha);
    /* block Givens rotation: QR of [rhobar; beta] (2nb x 2nb, last nb columns zero) */
    const int n2 = 2 * nb;
    memset(giv, 0, sizeof(double) * n2 * n2);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        giv[j * n2 + i] = rhobar[j * nb + i];
        giv[j * n2 + nb + i] = beta[j * nb + i];
      }
    orc_update_qrd(n2, 0, n2, giv, n2, tau);
    memset(ident, 0, sizeof(double) * n2 * n2);
    for (int i = 0; i < n2; ++i) ident[i * n2 + i] = 1.0;
    orc_mvmq(n2, n2, n2, giv, n2, tau, ident, n2, gq, n2);
    for (int c = 0; c < n2; ++c)
      for (int i = 0; i < n2; ++i) gr[c * n2 + i] = i <= c ? giv[c * n2 + i] : 0.0;
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        rho[j * nb + i] = gr[j * n2 + i];
        cc[j * nb + i] = gq[i * n2 + j];
        ss[j * nb + i] = gq[i * n2 + nb + j];
        ssbar[j * nb + i] = -gq[(i + nb) * n2 + j];
        ccbar[j * nb + i] = gq[(i + nb) * n2 + nb + j];
      }
    orc_ab(nb, ss, alpha, theta);     /* theta = ss alpha */
    orc_ab(nb, ccbar, alpha, rhobar); /* rhobar = ccbar alpha */
    orc_ab(nb, cc, phibar, phi);      /* phi = cc phibar */
    orc_ab(nb, ssbar, phibar, coef);  /* phibar = -ssbar phibar */
    for (int i = 0; i < nb2; ++i) phibar[i] = -coef[i];
    if (orc_inv_small(nb, rho, rinv)) { res.breakdown = 1; break; }
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        double aux = 0.0;
        for (int q = 0; q < nb; ++q) aux += rinv[q * nb + i] * phi[j * nb + q];
        coef[j * nb + i] = aux;
      }
    orc_daxpy_blk(n, nb, 1.0, coef, w, soli);
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) {
        double aux = 0.0;
        for (int q = 0; q < nb; ++q) aux += rinv[q * nb + i] * theta[j * nb + q];
        coef[j * nb + i] = aux;
      }
    memcpy(pw, v, vb);
    orc_daxpy_blk(n, nb, -1.0, coef, w, pw);
    memcpy(w, pw, vb);
    if (k % ichk == 0) {
      if (variant == 0) { orc_solvel(l, soli, t1, nb); orc_solveu(u, t1, xk, nb); }
      else if (variant == 1) memcpy(xk, soli, vb);
      else orc_solveu(u, soli, xk, nb);
      for (size_t i = 0; i < (size_t)n * nb; ++i) xk[i] += x0[i];
      orc_mvma(a, xk, r, nb);
      for (size_t i = 0; i < (size_t)n * nb; ++i) r[i] = b[i] - r[i];
      if (variant == 0) orc_blz_op(0, 1, a, l, u, nb, r, pw, t1, t2);
      conv = 0;
      double worst = 0.0;
      for (int c = 0; c < nb; ++c) {
        double s = 0.0, so = 0.0;
        for (int i = 0; i < n; ++i) s += r[(size_t)c * n + i] * r[(size_t)c * n + i];
        const double q = sqrt(s) / resi0[c];
        if (q > worst) worst = q;
        int ok = q < eps;
        if (variant == 0) {
          for (int i = 0; i < n; ++i) so += pw[(size_t)c * n + i] * pw[(size_t)c * n + i];
          ok = ok || sqrt(so) / resort0[c] < eps;
        }
        conv += ok;
      }
      res.resfin = worst;
      if (conv == nb) break;
    }
  }
  res.nitperf = k < niter ? k + 1 : niter;
  res.updates = conv == nb;
  if (variant == 0) { orc_solvel(l, soli, t1, nb); orc_solveu(u, t1, xk, nb); }
  else if (variant == 1) memcpy(xk, soli, vb);
  else orc_solveu(u, soli, xk, nb);
  for (size_t i = 0; i < (size_t)n * nb; ++i) sol_out[i] += xk[i];
  free(r); free(uu); free(v); free(w); free(qw); free(pw); free(t1); free(t2); free(soli); free(xk);
  return res;
}
