// ORACLE - test infrastructure only. C entry points around the reference's
// own solve-phase code (Bslib's CSMatrixRS / CSVector methods and genrcm),
// compiled together with the reference sources into oracle/_ref/ by
// oracle/Makefile.ref. Matrices are handed over in the arrays of the C ABI
// and placed into CSMatrixRS objects through the reference's public
// setters, the way AbstractSymmSolver builds its matrix (solvers.cpp:5383-5528).
// The library is loaded lazily (RTLD_LAZY): the parts of the compiled sources
// that would need MPI or METIS are never called. The IC2 / ILU2 entries run
// the reference's Ich2 / Ilu2 over its own LAPACK 3.0 (lapacksm/*.f and
// Bslib/lapackdm.f, compiled with flang into the same library).
#include <chrono>
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

// The same for the complex CSMatrixCS (smatrix.h:1499): a holds nza complex
// values interleaved (re, im), blocks in ja order.
CSMatrixCS make_cs(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a) {
  const int nzja = ia[nsup];
  int nza = 0, blamx = 0;
  for (int i = 0; i < nsup; ++i) {
    const int bi = sprnd[i + 1] - sprnd[i];
    if (bi > blamx) blamx = bi;
    for (int j = ia[i]; j < ia[i + 1]; ++j) nza += bi * (sprnd[ja[j] + 1] - sprnd[ja[j]]);
  }
  CSMatrixCS m(nsup, nzja, nza);
  m.SetM(sprnd[nsup]);
  m.SetN(sprnd[nsup]);
  m.SetNlist(nsup);
  m.SetNsupr(nsup);
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
  std::memcpy(m.GetA(), a, sizeof(double) * 2 * (size_t)nza);
  m.SetNsupc(nsup);
  m.SetNzja(nzja);
  m.SetNza(nza);
  m.SetNzatot(nza);
  return m;
}

void take_cs(const CSMatrixCS& m, int* nzja, int* nza, int* ia, int* ja, double* a) {
  int *psr, *psc, *pbsa, *plist, *pia, *pja;
  m.GetPartition(psr, psc, pbsa);
  m.GetSparsity(plist, pia, pja);
  const int nsup = m.GetNsupr();
  if (nzja) *nzja = pia[nsup];
  if (nza) *nza = m.GetNza();
  if (ia) std::memcpy(ia, pia, sizeof(int) * (nsup + 1));
  if (ja) std::memcpy(ja, pja, sizeof(int) * pia[nsup]);
  if (a) {
    int off = 0;
    for (int i = 0; i < nsup; ++i) {
      const int bi = psr[i + 1] - psr[i];
      for (int j = pia[i]; j < pia[i + 1]; ++j) {
        const int sz = bi * (psr[pja[j] + 1] - psr[pja[j]]);
        std::memcpy(a + 2 * (size_t)off, m.GetA() + pbsa[j], sizeof(double) * 2 * sz);
        off += sz;
      }
    }
  }
}

CSMatrixCS g_cfct[2];  // last complex factor: [0] L by columns, [1] U by rows

CSVector vec(int n, int nrhs, const double* x) {
  CSVector v(n, 1, nrhs);
  if (x) std::memcpy(v.GetVect(), x, sizeof(double) * n * nrhs);
  return v;
}

CSMatrixRS g_fct[2];  // last factor: [0] L by columns, [1] U by rows

void set_fct_params(CSlvParam& p, double tau1, double tau2, double theta, double sclmin, double pivmin,
                    double memory) {
  p.msglev = 0;
  p.tau1 = tau1;
  p.tau2 = tau2;
  p.theta = theta;
  p.sclmin = sclmin;
  p.pivmin = pivmin;
  p.memory = memory;
}

}  // namespace

extern "C" {

// CSMatrixRS::Ich2 (ich2rs.cpp:1075) of an upper-stored SPD matrix, as SolveSymm
// calls it (solvers.cpp:4545). Returns 0, or 1 when the reference threw (msg
// receives the text). The factor is kept for ref_fct_take(1, ...).
int ref_ich2(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, double tau1, double tau2,
             double theta, double memory, char* msg, int msglen) {
  CSMatrixRS am = make(nsup, sprnd, ia, ja, a);
  CSlvParam p;
  set_fct_params(p, tau1, tau2, theta, p.sclmin, p.pivmin, memory);
  std::ofstream fout("/dev/null");
  try {
    g_fct[1] = am.Ich2(fout, p);
    g_fct[0] = CSMatrixRS();
  } catch (const char* e) {
    std::strncpy(msg, e, msglen - 1);
    msg[msglen - 1] = 0;
    return 1;
  }
  return 0;
}

// CombineLU + CSMatrixRS::Ilu2 (ilu2rs.cpp:1458) of a full-storage matrix, as
// Solver(RS) calls them (solvers.cpp:1640-1660). Factors kept for ref_fct_take.
int ref_ilu2(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, double tau1, double tau2,
             double theta, double sclmin, double pivmin, double memory, char* msg, int msglen) {
  CSMatrixRS am = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS al, au, dummy;
  am.CombineLU(al, au);
  CSlvParam p;
  set_fct_params(p, tau1, tau2, theta, sclmin, pivmin, memory);
  std::ofstream fout("/dev/null");
  try {
    dummy.Ilu2(fout, p, al, au, g_fct[0], g_fct[1]);
  } catch (const char* e) {
    std::strncpy(msg, e, msglen - 1);
    msg[msglen - 1] = 0;
    return 1;
  }
  return 0;
}

// CSVector::BlockLanczosRight / Left / Central (lanczos.cpp:932, 1719, 2467)
// with ILU2 factors l (by columns) and u (by rows), as Solver(RS) calls them
// for ittype 3 / 4 / 2 (solvers.cpp:1798-1800); x0, b, sol are n x nrhs.
void ref_block_lanczos(int variant, int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                       const int* lia, const int* lja, const double* la, const int* uia, const int* uja,
                       const double* ua, int nrhs, const double* x0, const double* b, int niter, double eps, int ichk,
                       double* sol) {
  CSMatrixRS am = make(nsup, sprnd, ia, ja, a);
  CSMatrixRS lm = make(nsup, sprnd, lia, lja, la);
  CSMatrixRS um = make(nsup, sprnd, uia, uja, ua);
  const int n = sprnd[nsup];
  CSVector vx = vec(n, nrhs, x0), vb = vec(n, nrhs, b), out(n, 1, nrhs);
  out.SetSVect(0.0);
  CSlvParam p;
  p.msglev = 0;
  p.niter = niter;
  p.eps = eps;
  p.ichk = ichk;
  std::ofstream fout("/dev/null");
  if (variant == 0) out.BlockLanczosRight(fout, &am, &lm, &um, vx, vb, p);
  else if (variant == 1) out.BlockLanczosLeft(fout, &am, &lm, &um, vx, vb, p);
  else out.BlockLanczosCentral(fout, &am, &lm, &um, vx, vb, p);
  std::memcpy(sol, out.GetVect(), sizeof(double) * n * nrhs);
}

// CSMatrixRS::WriteMtrBin (smatrixrs.cpp:3210) of the matrix to path.
void ref_write_mtr_bin(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const char* path) {
  CSMatrixRS m = make(nsup, sprnd, ia, ja, a);
  std::ofstream f(path, std::ios::binary);
  m.WriteMtrBin(f);
}

// CSVector::WriteVectBin (svector.cpp:537) of an n x nrhs vector to path.
void ref_write_vect_bin(int n, int nrhs, const double* v, const char* path) {
  CSVector x = vec(n, nrhs, v);
  std::ofstream f(path, std::ios::binary);
  x.WriteVectBin(f);
}

// CSVector::ReadVectBin (svector.cpp:488) of path: sizes, then the values when v != null.
void ref_read_vect_bin(const char* path, int* n, int* nrhs, double* v) {
  std::ifstream f(path, std::ios::binary);
  CSVector x = CSVector::ReadVectBin(f);
  *n = x.GetNv();
  *nrhs = x.GetNrhs();
  if (v) std::memcpy(v, x.GetVect(), sizeof(double) * (size_t)x.GetNv() * x.GetNrhs());
}

// The last factor's part which (0 L, 1 U): sizes, then arrays when non-null.
void ref_fct_take(int which, int* nzja, int* nza, int* oia, int* oja, double* oa) {
  take(g_fct[which], nzja, nza, oia, oja, oa);
}

// Complex solve phase (mvmc.cpp:160, 1723, 2506): y = A x, y = L^-1 x,
// y = U^-1 x for CSMatrixCS, n x nrhs complex vectors interleaved (re, im).
void ref_cmvma(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* x, double* y,
               int nrhs) {
  CSMatrixCS m = make_cs(nsup, sprnd, ia, ja, a);
  const int n = sprnd[nsup];
  CSVectorC vx(n, 1, nrhs), vy(n, 1, nrhs);
  std::memcpy(vx.GetVect(), x, sizeof(double) * 2 * (size_t)n * nrhs);
  m.MvmA(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * 2 * (size_t)n * nrhs);
}

void ref_csolve(int which, int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                const double* x, double* y, int nrhs) {
  CSMatrixCS m = make_cs(nsup, sprnd, ia, ja, a);
  const int n = sprnd[nsup];
  CSVectorC vx(n, 1, nrhs), vy(n, 1, nrhs);
  std::memcpy(vx.GetVect(), x, sizeof(double) * 2 * (size_t)n * nrhs);
  if (which == 0) m.SolveL(vx, vy);
  else m.SolveU(vx, vy);
  std::memcpy(y, vy.GetVect(), sizeof(double) * 2 * (size_t)n * nrhs);
}

// CombineLU + CSMatrixCS::Ilu2 (ilu2cs.cpp) of a complex matrix, the factors
// kept for ref_cfct_take. Returns 0, or 1 when the reference threw.
int ref_cilu2(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, double tau1, double tau2,
              double theta, double sclmin, double pivmin, double memory, int fcttyp, char* msg, int msglen) {
  CSMatrixCS am = make_cs(nsup, sprnd, ia, ja, a);
  CSMatrixCS al, au, dummy;
  am.CombineLU(al, au);
  CSlvParam p;
  set_fct_params(p, tau1, tau2, theta, sclmin, pivmin, memory);
  p.fcttyp = fcttyp;
  std::ofstream fout("/dev/null");
  try {
    dummy.Ilu2(fout, p, al, au, g_cfct[0], g_cfct[1]);
  } catch (const char* e) {
    std::strncpy(msg, e, msglen - 1);
    msg[msglen - 1] = 0;
    return 1;
  }
  return 0;
}

void ref_cfct_take(int which, int* nzja, int* nza, int* oia, int* oja, double* oa) {
  take_cs(g_cfct[which], nzja, nza, oia, oja, oa);
}

// The reference's complex BlockGMRES (gmres.cpp:1044-1466) of one
// right-hand side with the factors of the last ref_cilu2 call: niter steps
// at most, stopped at ||r|| < eps ||r0||; x holds x0 on entry and the
// solution on return. Returns the seconds the call took (host clock).
double ref_cblock_gmres(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, const double* b,
                        double* x, int niter, double eps) {
  CSMatrixCS am = make_cs(nsup, sprnd, ia, ja, a);
  const int n = sprnd[nsup];
  CSVectorC vb(n, 1, 1), vx(n, 1, 1), sol(n, 1, 1);
  std::memcpy(vb.GetVect(), b, sizeof(double) * 2 * (size_t)n);
  std::memcpy(vx.GetVect(), x, sizeof(double) * 2 * (size_t)n);
  CSlvParam p;
  p.msglev = 0;
  p.niter = niter;
  p.eps = eps;
  p.ichk = niter + 1;
  std::ofstream fout("/dev/null");
  const auto t0 = std::chrono::steady_clock::now();
  sol.BlockGMRES(fout, &am, &g_cfct[0], &g_cfct[1], vx, vb, p);
  const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::memcpy(x, sol.GetVect(), sizeof(double) * 2 * (size_t)n);
  return sec;
}

// zgesvd_ as the complex ILU2 calls it (ilu2cs.cpp:132, 725): through
// lapackd.cpp:944 -> lapackdm.f -> ZGESVD_SM, values interleaved re/im.
void ref_zgesvd(int n, double* a, double* s, double* u, double* vt, double* work, int lwork, double* rwork,
                int* info) {
  zgesvd_("A", "A", &n, &n, reinterpret_cast<dcmplx*>(a), &n, s, reinterpret_cast<dcmplx*>(u), &n,
          reinterpret_cast<dcmplx*>(vt), &n, reinterpret_cast<dcmplx*>(work), &lwork, rwork, info);
}

// Index arrays for a CIAJA callback: the reference releases them with delete[]
// (solvers.cpp:7388-7389).
int* ref_alloc_ints(int n) { return new int[n > 0 ? n : 1]; }

// The reference's complex AbstractSolver (solvers.cpp:7103-7398) on the
// caller's callbacks: assembly, CSVectorC::Solver (solvers.cpp:2588-2830:
// ordering by ordtype, CombineLU + Ilu2, block GMRES, inverse ordering) and
// CSTORESOLC. Returns 0, or 1 when the reference threw.
int ref_abstract_solver_c(void* obj, int ordtype, int fcttyp, double tau1, double tau2, double theta, double sclmin,
                          double pivmin, double memory, int niter, double eps, CNSUP fnsup, CSPRNDS fsprnds,
                          CIAJA fiaja, CBLOCKROWC fblockrow, CRHSC frhs, CINIGUESSC finiguess,
                          CSTORESOLC fstoresol, char* msg, int msglen) {
  CSlvParam p;
  set_fct_params(p, tau1, tau2, theta, sclmin, pivmin, memory);
  p.ordtype = ordtype;
  p.fcttyp = fcttyp;
  p.niter = niter;
  p.eps = eps;
  std::ofstream fout("/dev/null");
  try {
    AbstractSolver(obj, fout, p, fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol);
  } catch (const char* e) {
    std::strncpy(msg, e, msglen - 1);
    msg[msglen - 1] = 0;
    return 1;
  }
  return 0;
}

// dsyev_ / dgesvd_ as the factorizations call them (lapackd.cpp:865-905), for
// pinning the dense kernels on their own: job strings are the literals the
// callers pass, so the pointer comparison inside lapackd.cpp sees the same
// (merged) literals as in ich2rs.cpp / ilu2rs.cpp.
void ref_dsyev(int n, double* a, double* w, double* work, int lwork, int* info) {
  dsyev_("V", "U", &n, a, &n, w, work, &lwork, info);
}
void ref_dgesvd(int n, double* a, double* s, double* u, double* vt, double* work, int lwork, int* info) {
  dgesvd_("A", "A", &n, &n, a, &n, s, u, &n, vt, &n, work, &lwork, info);
}

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

// CSMatrix::FindSeparator (smatrix.cpp:4355) on an n-row pattern (unit blocks).
void ref_find_separator(int n, const int* ia, const int* ja, int* ind1, int* ind2) {
  int* sp = new int[n + 1];
  for (int i = 0; i <= n; ++i) sp[i] = i;
  double* a = new double[ia[n] > 0 ? ia[n] : 1]();
  CSMatrixRS m = make(n, sp, ia, ja, a);
  int i1 = 0, i2 = 0;
  m.FindSeparator(i1, i2);
  *ind1 = i1;
  *ind2 = i2;
  delete[] sp;
  delete[] a;
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
