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
 * solve agreement, monotone FGMRES residuals).
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
