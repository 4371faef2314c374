// Second-order incomplete block Cholesky (IC2) on the host.
//
// Restates CSMatrixRS::Ich2 (ich2rs.cpp:1075-1378) and its helpers
// Ich2AllocScale (:26-245), Ich2InitRow (:276-363), Ich2InitMask (:369-411),
// Ich2UpdateRow (:416-462), Ich2FiltrRow (:467-527), Ich2PivScaleRow
// (:532-644), Ich2FreeR (:649-743), Ich2StoreRow (:748-817), Ich2StoreScaleU
// (:822-940), with the integer machinery of ich2r.cpp:27-55 (Ich2AllocInt)
// and :296-327 (Ich2UpdateTransp).
//
// The row-by-row left-looking sweep, its linked lists of pending rows, the
// order in which every update is subtracted and the arithmetic of every
// scaling product follow the reference step for step, because the factor's
// values (and the drop decisions taken on them) depend on that order.
#include "bcm_host.hpp"
#include "block_ops.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace bcm {

namespace {

struct Ic2State {
  const BlockCsr& A;
  int nsup;
  std::vector<int64_t> bsdia;   // offset of the bla_i^2 dense block of row i
  std::vector<double> scl, sclinv, diagg;
  std::vector<double> fmask;    // dense row buffer: block (irow, j) at sprnd[j]*bla_irow
  // factor storage (rows appended in order; R entries carry negative columns)
  std::vector<int> ig, jg;
  std::vector<int64_t> addrg;
  std::vector<double> g;
  explicit Ic2State(const BlockCsr& a) : A(a), nsup(a.nsup) {}
  int bla(int i) const { return A.sprnd[i + 1] - A.sprnd[i]; }
};

}  // namespace

BlockCsr ich2(const BlockCsr& A, const SlvParam& prm, FactorStats* st) {
  auto t0 = std::chrono::steady_clock::now();
  A.validate(true);
  for (int i = 0; i < A.nsup; ++i)
    for (int j = A.ia[i]; j < A.ia[i + 1]; ++j)
      if (A.ja[j] < i) throw Error(kErrArg, "ich2: matrix must be stored upper-triangular");
  Ic2State S(A);
  const int nsup = A.nsup;
  const int n = A.n();
  const int blamx = std::max(1, [&] { int m = 0; for (int i = 0; i < nsup; ++i) m = std::max(m, A.bla(i)); return m; }());

  // ---- scaling of the diagonal blocks (Ich2AllocScale) ----
  S.bsdia.resize(nsup + 1);
  int64_t nzd = 0;
  for (int i = 0; i < nsup; ++i) { S.bsdia[i] = nzd; nzd += (int64_t)S.bla(i) * S.bla(i); }
  S.bsdia[nsup] = nzd;
  S.scl.resize(nzd);
  S.sclinv.resize(nzd);
  S.diagg.resize(nzd);
  std::vector<double> aloc(blamx * blamx), vloc(blamx * blamx), eig(blamx), work(blamx),
      delem(blamx * blamx), lelem(blamx * blamx);
  for (int i = 0; i < nsup; ++i) {
    const int b = S.bla(i);
    const double* aii = &A.a[A.bsa[A.ia[i]]];
    std::memcpy(aloc.data(), aii, sizeof(double) * b * b);
    sym_eig(b, aloc.data(), eig.data());
    for (int k = 0; k < b; ++k)
      if (eig[k] < 0.0) throw Error(kErrNotSpd, " Negative eigenvalue when scaling ");
    for (int k = 0; k < b; ++k) eig[k] = std::sqrt(eig[k]);
    double* scl = &S.scl[S.bsdia[i]];
    double* sci = &S.sclinv[S.bsdia[i]];
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) sci[c * b + r] = aloc[c * b + r] * eig[c];
    for (int k = 0; k < b; ++k) eig[k] = 1.0 / eig[k];
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) scl[c * b + r] = aloc[r * b + c] * eig[r];
    // diagg = scl * A_ii * scl^T
    left_mul(b, b, scl, aii, aloc.data());
    right_mul_t(b, b, aloc.data(), scl, vloc.data());
    std::memcpy(&S.diagg[S.bsdia[i]], vloc.data(), sizeof(double) * b * b);
  }

  // ---- integer work arrays (Ich2AllocInt) ----
  int icycle = -1;
  std::vector<int> imask(nsup, -1), lstloc(nsup), iv(nsup, -1), madj(nsup, -1), ibegm(nsup, -1),
      lstupd(nsup), indlst(nsup);
  S.fmask.assign((size_t)n * blamx, 0.0);
  double* fmask = S.fmask.data();

  // ---- factor storage ----
  const int64_t cap_vals = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)A.nza()) : -1;
  const int64_t cap_blocks = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)A.nzja()) : -1;
  S.ig.assign(nsup + 1, 0);
  S.jg.clear();
  S.addrg.clear();
  S.g.clear();
  int ibegfr = 0;
  int64_t nzjrl = 0, nzrtot = 0, nzju = 0, nzjr = 0;
  int64_t compact_mark = std::max<int64_t>(A.nza(), 1 << 20);
  if (cap_vals < 0) {  // storage stays below twice the compaction mark plus one row: reserve it once
    S.g.reserve((size_t)(2 * compact_mark) + (size_t)blamx * blamx * 64);
    S.jg.reserve((size_t)2 * A.nzja() + 64);
    S.addrg.reserve((size_t)2 * A.nzja() + 64);
  }

  // Removes second-order blocks whose column lies before `irow`: they can no
  // longer take part in any update (Ich2FreeR). Values are untouched.
  auto free_r = [&](int irow) {
    int64_t nzjg = S.ig[ibegfr];
    int64_t nznew = ibegfr != 0 ? S.addrg[S.ig[ibegfr]] : 0;
    int64_t beg = S.ig[ibegfr];
    int ibgfrn = ibegfr;
    nzjrl = 0;
    for (int i = ibegfr; i < irow; ++i) {
      const int bi = S.bla(i);
      for (int64_t j = beg; j < S.ig[i + 1]; ++j) {
        int col = S.jg[j];
        if (col < 0) {
          if (-col < irow) continue;
          ++nzjrl;
          col = -col;
        }
        const int64_t sz = (int64_t)bi * S.bla(col);
        if (nzjg != j) {
          S.jg[nzjg] = S.jg[j];
          std::memmove(&S.g[nznew], &S.g[S.addrg[j]], sizeof(double) * sz);
        }
        S.addrg[nzjg] = nznew;
        nznew += sz;
        ++nzjg;
      }
      beg = S.ig[i + 1];
      S.ig[i + 1] = (int)nzjg;
      if (nzjrl == 0 && ibgfrn == i) ibgfrn = i + 1;
      iv[i] += (int)(nzjg - beg);
    }
    ibegfr = ibgfrn;
    S.jg.resize(nzjg);
    S.addrg.resize(nzjg);
    S.g.resize(nznew);
  };

  for (int irow = 0; irow < nsup; ++irow) {
    ++icycle;
    const int bi = S.bla(irow);
    const int ibsv = A.sprnd[irow];

    // ---- init the scaled row (Ich2InitRow) ----
    int nlist1 = 0;
    lstloc[nlist1++] = irow;
    imask[irow] = icycle;
    std::memcpy(fmask + (int64_t)ibsv * bi, &S.diagg[S.bsdia[irow]], sizeof(double) * bi * bi);
    const double* scli = &S.scl[S.bsdia[irow]];
    for (int j = A.ia[irow] + 1; j < A.ia[irow + 1]; ++j) {
      const int jj = A.ja[j];
      lstloc[nlist1++] = jj;
      imask[jj] = icycle;
      const int bj = S.bla(jj);
      const double* aij = &A.a[A.bsa[j]];
      const double* sclj = &S.scl[S.bsdia[jj]];
      // aloc = scl_i * A_ij ; vloc = aloc * scl_j^T
      left_mul(bi, bj, scli, aij, aloc.data());
      right_mul_t(bi, bj, aloc.data(), sclj, vloc.data());
      std::memcpy(fmask + (int64_t)A.sprnd[jj] * bi, vloc.data(), sizeof(double) * bi * bj);
    }

    // ---- left-looking updates by the pending rows ----
    int irwprv = ibegm[irow];
    while (irwprv != -1) {
      const int64_t jpos = iv[irwprv];
      const int icolmn = S.jg[jpos];
      const int bp = S.bla(irwprv);
      std::memcpy(lelem.data(), &S.g[S.addrg[jpos]], sizeof(double) * bp * bi);
      const bool elmisr = icolmn < 0;
      // Ich2InitMask: register the columns this row contributes to.
      int nupd = 0;
      for (int64_t k = iv[irwprv]; k < S.ig[irwprv + 1]; ++k) {
        int jc = S.jg[k];
        if (elmisr && jc < 0) continue;  // second-order x second-order is skipped
        if (jc < 0) jc = -jc;
        if (imask[jc] != icycle) {
          imask[jc] = icycle;
          lstloc[nlist1++] = jc;
          std::memset(fmask + (int64_t)A.sprnd[jc] * bi, 0, sizeof(double) * bi * S.bla(jc));
        }
        lstupd[nupd] = jc;
        indlst[nupd] = (int)(k - iv[irwprv]);
        ++nupd;
      }
      // Ich2UpdateRow: F(irow, col) -= G(p, irow)^T G(p, col), one product at a time.
      for (int u = 0; u < nupd; ++u) {
        const int col = lstupd[u];
        const int bj = S.bla(col);
        double* f = fmask + (int64_t)A.sprnd[col] * bi;
        const double* gb = &S.g[S.addrg[iv[irwprv] + indlst[u]]];
        sub_tprod(bi, bj, bp, lelem.data(), gb, f);
      }
      irwprv = madj[irwprv];
    }

    // ---- move the pending rows on to their next column (Ich2UpdateTransp) ----
    {
      int p = ibegm[irow];
      while (p != -1) {
        const int pnext = madj[p];
        if (iv[p] >= S.ig[p + 1] - 1) {
          madj[p] = -1;
        } else {
          int col = S.jg[iv[p] + 1];
          if (col < 0) col = -col;
          madj[p] = ibegm[col];
          ibegm[col] = p;
        }
        iv[p]++;
        p = pnext;
      }
    }

    // ---- drop small blocks with diagonal compensation (Ich2FiltrRow) ----
    {
      int nlist2 = 0;
      double* fdiag = fmask + (int64_t)ibsv * bi;
      for (int i = 1; i < nlist1; ++i) {
        const int j = lstloc[i];
        const int bj = S.bla(j);
        const double* f = fmask + (int64_t)A.sprnd[j] * bi;
        double amax = 0.0;
        for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(f[k]));
        if (amax >= prm.tau2) {
          lstupd[nlist2++] = j;
        } else {
          double* dj = &S.diagg[S.bsdia[j]];
          for (int r = 0; r < bi; ++r)
            for (int c = 0; c < bj; ++c) {
              double v = std::fabs(f[c * bi + r]) * prm.theta;
              fdiag[r * bi + r] += v;
              dj[c * bj + c] += v;
            }
        }
      }
      nlist1 = nlist2;
    }

    // ---- pivot: inverse square root of the diagonal block, scale the row (Ich2PivScaleRow) ----
    {
      double* fdiag = fmask + (int64_t)ibsv * bi;
      std::memcpy(aloc.data(), fdiag, sizeof(double) * bi * bi);
      sym_eig(bi, aloc.data(), eig.data());
      // descending order of the eigenpairs
      for (int k = 0; k < bi; ++k)
        for (int m = 0; m < bi; ++m) delem[(bi - k - 1) * bi + m] = aloc[k * bi + m];
      std::memcpy(aloc.data(), delem.data(), sizeof(double) * bi * bi);
      for (int k = 0; k < bi; ++k) work[bi - k - 1] = eig[k];
      for (int k = 0; k < bi; ++k) eig[k] = work[k];
      for (int k = 0; k < bi; ++k) {
        if (eig[k] < 0.0) throw Error(kErrNotSpd, " Negative pivot in ICH2 ");
        eig[k] = std::sqrt(eig[k]);
        eig[k] = 1.0 / eig[k];
      }
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bi; ++c) delem[c * bi + r] = aloc[r * bi + c] * eig[r];
      for (int i = 0; i < nlist1; ++i) {
        const int j = lstupd[i];
        const int bj = S.bla(j);
        double* f = fmask + (int64_t)A.sprnd[j] * bi;
        left_mul(bi, bj, delem.data(), f, aloc.data());
        std::memcpy(f, aloc.data(), sizeof(double) * bi * bj);
      }
    }

    // ---- workspace bound ----
    int64_t nzloc = (int64_t)bi * bi;
    for (int i = 0; i < nlist1; ++i) nzloc += (int64_t)bi * S.bla(lstupd[i]);
    int64_t nzjg = (int64_t)S.jg.size();
    int64_t nznew = (int64_t)S.g.size();
    if (cap_vals > 0) {
      if (nzjg + nlist1 + 1 > cap_blocks || nznew + nzloc > cap_vals) {
        free_r(irow);
        nzjg = (int64_t)S.jg.size();
        nznew = (int64_t)S.g.size();
      }
      if (nzjg + nlist1 + 1 > cap_blocks || nznew + nzloc > cap_vals)
        throw Error(kErrWorkspace, " Insufficient workspace in Ich2");
    } else if (nznew + nzloc > 2 * compact_mark) {
      free_r(irow);
      nzjg = (int64_t)S.jg.size();
      nznew = (int64_t)S.g.size();
      compact_mark = std::max(compact_mark, nznew);
    }

    std::sort(lstupd.begin(), lstupd.begin() + nlist1);
    if (nlist1 != 0) {
      iv[irow] = (int)(nzjg + 1);
      madj[irow] = ibegm[lstupd[0]];
      ibegm[lstupd[0]] = irow;
    }

    // ---- store the row (Ich2StoreRow): inverse pivot transposed, then the
    // first-order (>= tau1) and second-order blocks ----
    S.jg.push_back(irow);
    S.addrg.push_back(nznew);
    S.g.resize(nznew + nzloc);
    double* gd = &S.g[nznew];
    for (int r = 0; r < bi; ++r)
      for (int c = 0; c < bi; ++c) gd[c * bi + r] = delem[r * bi + c];
    nznew += (int64_t)bi * bi;
    ++nzju;
    for (int i = 0; i < nlist1; ++i) {
      const int j = lstupd[i];
      const int bj = S.bla(j);
      const double* f = fmask + (int64_t)A.sprnd[j] * bi;
      double amax = 0.0;
      for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(f[k]));
      S.addrg.push_back(nznew);
      if (amax >= prm.tau1) {
        S.jg.push_back(j);
        ++nzju;
      } else {
        S.jg.push_back(-j);
        ++nzjr;
        ++nzjrl;
        nzrtot += (int64_t)bi * bj;
      }
      std::memcpy(&S.g[nznew], f, sizeof(double) * bi * bj);
      nznew += (int64_t)bi * bj;
    }
    if ((int64_t)S.jg.size() > INT32_MAX) throw Error(kErrWorkspace, " Insufficient workspace in Ich2");
    S.ig[irow + 1] = (int)S.jg.size();
  }

  // ---- unscale and assemble U (Ich2StoreScaleU) ----
  BlockCsr U;
  U.nsup = nsup;
  U.sprnd = A.sprnd;
  U.ia.assign(nsup + 1, 0);
  int64_t nz = 0;
  for (int i = 0; i < nsup; ++i)
    for (int64_t j = S.ig[i]; j < S.ig[i + 1]; ++j)
      if (S.jg[j] >= 0) nz += (int64_t)S.bla(i) * S.bla(S.jg[j]);
  U.a.resize(nz);
  nz = 0;
  for (int i = 0; i < nsup; ++i) {
    const int bi = S.bla(i);
    const double* scli = &S.scl[S.bsdia[i]];
    for (int64_t j = S.ig[i]; j < S.ig[i + 1]; ++j) {
      const int jj = S.jg[j];
      if (jj < 0) continue;
      const int bj = S.bla(jj);
      const double* gb = &S.g[S.addrg[j]];
      if (jj == i) {
        // D_i = scl_i^T * G_ii
        left_mul_t(bi, bi, scli, gb, aloc.data());
      } else {
        // U_ij = G_ij * sclinv_j^T
        const double* scij = &S.sclinv[S.bsdia[jj]];
        right_mul_t(bi, bj, gb, scij, aloc.data());
      }
      std::memcpy(&U.a[nz], aloc.data(), sizeof(double) * bi * bj);
      U.ja.push_back(jj);
      U.bsa.push_back(nz);
      nz += (int64_t)bi * bj;
    }
    U.ia[i + 1] = (int)U.ja.size();
  }
  U.set_blamx();
  if (st) {
    st->nzu = nzju;
    st->nzr = nzjr;
    st->densu = (double)U.nza() / (double)A.nza();
    st->densr = (double)nzrtot / (double)A.nza();
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return U;
}

}  // namespace bcm
