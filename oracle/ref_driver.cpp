// ORACLE - test infrastructure only. C entry points around the reference's
// own solve-phase code (Bslib's CSMatrixRS / CSVector methods and genrcm),
// compiled together with the reference sources into oracle/_ref/ by
// oracle/Makefile.ref. Matrices are handed over in the arrays of the C ABI
// and placed into CSMatrixRS objects through the reference's public
// setters, the way AbstractSymmSolver builds its matrix (solvers.cpp:5383-5528).
// The library is loaded lazily (RTLD_LAZY): the parts of the compiled sources
// that would need MPI, METIS or LAPACK are never called.
#include <cstring>
#include <fstream>

#include "smatrix.h"
#include "slvparam.h"
#include "svector.h"

extern void genrcm(int n, int* xadj, int* iadj, int* perm, int* xls, int* mask);

namespace {

CSMatrixRS make(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a) {
  const int nzja = ia[nsup];
  int nza = 0, blamx = 0;
  for (int i = 0; i < nsup; ++i) {
    const int bi = sprnd[i + 1] - sprnd[i];
    if (bi > blamx) blamx = bi;
    for (int j = ia[i]; j < ia[i + 1]; ++j) nza += bi * (sprnd[ja[j] + 1] - sprnd[ja[j]]);
  }
  CSMatrixRS m(nsup, nzja, nza);
  m.SetM(sprnd[nsup]);
  m.SetN(sprnd[nsup]);
  m.SetNlist(nsup);
  m.SetNsupr(nsup);
  m.SetNsupcBase(nsup);
  m.SetBlamx(blamx);
  int *psr, *psc, *pbsa, *plist, *pia, *pja;
  m.GetPartition(psr, psc, pbsa);
  m.GetSparsity(plist, pia, pja);
  for (int i = 0; i <= nsup; ++i) psr[i] = psc[i] = sprnd[i];
  for (int i = 0; i < nsup; ++i) plist[i] = i;
  for (int i = 0; i <= nsup; ++i) pia[i] = ia[i];
  int off = 0;
  for (int i = 0; i < nsup; ++i) {
    const int bi = sprnd[i + 1] - sprnd[i];
    for (int j = ia[i]; j < ia[i + 1]; ++j) {
      pja[j] = ja[j];
      pbsa[j] = off;
      off += bi * (sprnd[ja[j] + 1] - sprnd[ja[j]]);
    }
  }
  std::memcpy(m.GetA(), a, sizeof(double) * nza);
  m.SetNzja(nzja);
  m.SetNza(nza);
  m.SetNzatot(nza);
  return m;
}

// Copies a CSMatrixRS out: sizes first (out_nzja/out_nza), arrays when non-null.
void take(const CSMatrixRS& m, int* nzja, int* nza, int* ia, int* ja, double* a) {
  int *psr, *psc, *pbsa, *plist, *pia, *pja;
  m.GetPartition(psr, psc, pbsa);
  m.GetSparsity(plist, pia, pja);
  const int nsup = m.GetNsupr();
  if (nzja) *nzja = pia[nsup];
  if (nza) *nza = m.GetNza();
  if (ia) std::memcpy(ia, pia, sizeof(int) * (nsup + 1));
  if (ja) std::memcpy(ja, pja, sizeof(int) * pia[nsup]);
  if (a) {
    // blocks in ja order (bsa may not be contiguous in general)
    int off = 0;
    for (int i = 0; i < nsup; ++i) {
      const int bi = psr[i + 1] - psr[i];
      for (int j = pia[i]; j < pia[i + 1]; ++j) {
        const int sz = bi * (psr[pja[j] + 1] - psr[pja[j]]);
        std::memcpy(a + off, m.GetA() + pbsa[j], sizeof(double) * sz);
        off += sz;
      }
    }
  }
}

CSVector vec(int n, int nrhs, const double* x) {
  CSVector v(n, 1, nrhs);
  if (x) std::memcpy(v.GetVect(), x, sizeof(double) * n * nrhs);
  return v;
}

}  // namespace

extern "C" {

void ref_mvma(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x, double* y,
              int nrhs) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSVector vx = vec(sprnd[nsup], nrhs, x), vy = vec(sprnd[nsup], nrhs, nullptr);
  m.MvmA(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * sprnd[nsup] * nrhs);
}

void ref_mvmat(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x,
               double* y, int nrhs) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSVector vx = vec(sprnd[nsup], nrhs, x), vy = vec(sprnd[nsup], nrhs, nullptr);
  m.MvmAt(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * sprnd[nsup] * nrhs);
}

void ref_mvmasymm(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x,
                  double* y) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSVector vx = vec(sprnd[nsup], 1, x), vy = vec(sprnd[nsup], 1, nullptr);
  m.MvmASymm(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * sprnd[nsup]);
}

void ref_solvel(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x,
                double* y, int nrhs) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSVector vx = vec(sprnd[nsup], nrhs, x), vy = vec(sprnd[nsup], nrhs, nullptr);
  m.SolveL(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * sprnd[nsup] * nrhs);
}

void ref_solveu(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x,
                double* y, int nrhs) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSVector vx = vec(sprnd[nsup], nrhs, x), vy = vec(sprnd[nsup], nrhs, nullptr);
  m.SolveU(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * sprnd[nsup] * nrhs);
}

double ref_scprod(int n, const double* x, const double* y) {
  CSVector vx = vec(n, 1, x), vy = vec(n, 1, y);
  return vx.ScProdSVect(vy);
}

// CSVector::Cg (cg.cpp:31) on an upper-stored matrix with the IC2 factor u.
void ref_cg(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const int* uia,
            const int* uja, const double* ua, const double* x0, const double* b, int niter, double eps, int ittype,
            double* sol, int* nitperf, double* resfin) {
  CSMatrixRS am = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS um = make(nsup, sprnd, uia, uja, ua);
  const int n = sprnd[nsup];
  CSVector vx = vec(n, 1, x0), vb = vec(n, 1, b), norm = vec(n, 1, nullptr), out(n);
  norm.SetSVect(1.0);
  CSlvParam p;
  p.msglev = 0;
  p.niter = niter;
  p.eps = eps;
  p.ittype = ittype;
  std::ofstream fout("/dev/null");
  // istore = 0: the distributed factor argument is never touched (cg.cpp:79-92, :141-146)
  out.Cg(fout, &am, 0, (CGSMatrix*)0, &um, vx, vb, p, norm);
  std::memcpy(sol, out.GetVect(), sizeof(double) * n);
  *nitperf = p.nitperf;
  *resfin = p.resfin;
}

// CSVector::Lanczos serial form (lanczos.cpp:34) with ILU2 factors l (by columns) and u (by rows).
void ref_lanczos(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const int* lia,
                 const int* lja, const double* la, const int* uia, const int* uja, const double* ua,
                 const double* x0, const double* b, int niter, double eps, int sttype, double* sol, int* nitperf,
                 double* resfin) {
  CSMatrixRS am = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS lm = make(nsup, sprnd, lia, lja, la);
  CSMatrixRS um = make(nsup, sprnd, uia, uja, ua);
  const int n = sprnd[nsup];
  CSVector vx = vec(n, 1, x0), vb = vec(n, 1, b), norm = vec(n, 1, nullptr), out(n);
  norm.SetSVect(1.0);
  CSlvParam p;
  p.msglev = 0;
  p.niter = niter;
  p.eps = eps;
  p.sttype = sttype;
  std::ofstream fout("/dev/null");
  out.Lanczos(fout, &am, &lm, &um, vx, vb, p, norm);
  std::memcpy(sol, out.GetVect(), sizeof(double) * n);
  *nitperf = p.nitperf;
  *resfin = p.resfin;
}

// genrcm (rcm.cpp:216) on a 1-based adjacency; perm is 1-based on return.
void ref_genrcm(int n, int* xadj1, int* adj1, int* perm1) {
  int* xls = new int[n + 1];
  int* mask = new int[n + 1];
  genrcm(n, xadj1, adj1, perm1, xls, mask);
  delete[] xls;
  delete[] mask;
}

// OrderPrfMtr(1) (ordmatr.cpp:28): order[old] = new.
void ref_order_prf(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, int* order) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  m.OrderPrfMtr(1, order);
}

// OrdMtrSymm (smatrixrs.cpp:456) / OrdMtr (:345): two calls, sizes then arrays.
void ref_ordmtr(int symm, int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                const int* order, int* nzja, int* nza, int* oia, int* oja, double* oa) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS r = symm ? m.OrdMtrSymm(order) : m.OrdMtr(order);
  take(r, nzja, nza, oia, oja, oa);
}

// CombineLU (smatrixrs.cpp:1645): which 0 -> L by columns, 1 -> U by rows.
void ref_combinelu(int which, int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, int* nzja,
                   int* nza, int* oia, int* oja, double* oa) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS l, u;
  m.CombineLU(l, u);
  take(which == 0 ? l : u, nzja, nza, oia, oja, oa);
}

}  // extern "C"
