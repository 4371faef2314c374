// Second-order incomplete block LU (ILU2) on the host.
//
// Restates CSMatrixRS::Ilu2 (ilu2rs.cpp:1458-1817) and its helpers
// Ilu2AllocScale (:25-311, SVD scaling with the sclmin floor), Ilu2InitRow
// (:377-508), Ilu2InitMask (:514-560), Ilu2FiltrRow (:616-688),
// Ilu2PivScaleRow (:694-849, SVD pivots with the pivmin floor), Ilu2FreeR
// (:854-961), Ilu2StoreRow (:1189-1283); the two-mask update uses
// Ich2UpdateRow twice (:1603-1606) and the result is unscaled by
// Ich2StoreScaleU for L and for U (:1757-1764).
//
// Input: L stored by columns (the transposed lower blocks) and U by rows with
// identical patterns (combine_lu). Output: L by columns and U by rows, both
// with inverted diagonal blocks, A ~ L U.
#include "bcm_host.hpp"
#include "block_ops.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace bcm {

namespace {

// Two-sided scaling factors of one diagonal block from a = U diag(s) V^T:
// scll = s^-1/2 U^T, sclu = s^-1/2 V^T, sclinvl = U s^1/2, sclinvu = V s^1/2.
void svd_scaling(int b, const double* a, double floor_, double* scll, double* sclu, double* sclinvl,
                 double* sclinvu, double* u, double* vt, double* s) {
  if (b > 1) {
    svd_square(b, a, s, u, vt);
  } else {
    u[0] = 1.0;
    if (a[0] >= 0.0) { vt[0] = 1.0; s[0] = a[0]; }
    else { vt[0] = -1.0; s[0] = -a[0]; }
  }
  for (int k = 0; k < b; ++k)
    if (s[k] < floor_) s[k] = floor_;
  for (int k = 0; k < b; ++k) s[k] = std::sqrt(s[k]);
  for (int r = 0; r < b; ++r)
    for (int c = 0; c < b; ++c) sclinvl[c * b + r] = u[c * b + r] * s[c];
  for (int r = 0; r < b; ++r)
    for (int c = 0; c < b; ++c) sclinvu[c * b + r] = vt[r * b + c] * s[c];
  for (int k = 0; k < b; ++k) s[k] = 1.0 / s[k];
  for (int r = 0; r < b; ++r)
    for (int c = 0; c < b; ++c) scll[c * b + r] = u[r * b + c] * s[r];
  for (int r = 0; r < b; ++r)
    for (int c = 0; c < b; ++c) sclu[c * b + r] = vt[c * b + r] * s[r];
}

}  // namespace

void ilu2(const BlockCsr& AL, const BlockCsr& AU, const SlvParam& prm, BlockCsr& Lout, BlockCsr& Uout,
          FactorStats* st) {
  auto t0 = std::chrono::steady_clock::now();
  AL.validate(true);
  AU.validate(true);
  if (AL.nsup != AU.nsup || AL.ja != AU.ja || AL.sprnd != AU.sprnd)
    throw Error(kErrArg, "ilu2: L and U parts must share one pattern");
  const int nsup = AL.nsup;
  const int n = AL.n();
  auto bla = [&](int i) { return AL.sprnd[i + 1] - AL.sprnd[i]; };
  int blamx = 1;
  for (int i = 0; i < nsup; ++i) blamx = std::max(blamx, bla(i));

  std::vector<int64_t> bsdia(nsup + 1);
  int64_t nzd = 0;
  for (int i = 0; i < nsup; ++i) { bsdia[i] = nzd; nzd += (int64_t)bla(i) * bla(i); }
  bsdia[nsup] = nzd;
  std::vector<double> scll(nzd), sclu(nzd), sclinvl(nzd), sclinvu(nzd), diagg(nzd);
  const int bb = blamx * blamx;
  std::vector<double> aloc(bb), vloc(bb), uloc(bb), vtloc(bb), eig(blamx), deleml(bb), delemu(bb),
      lelem(bb), uelem(bb);

  // ---- scaling (Ilu2AllocScale) ----
  for (int i = 0; i < nsup; ++i) {
    const int b = bla(i);
    const double* aii = &AU.a[AU.bsa[AU.ia[i]]];
    svd_scaling(b, aii, prm.sclmin, &scll[bsdia[i]], &sclu[bsdia[i]], &sclinvl[bsdia[i]],
                &sclinvu[bsdia[i]], uloc.data(), vtloc.data(), eig.data());
    left_mul(b, b, &scll[bsdia[i]], aii, aloc.data());
    right_mul_t(b, b, aloc.data(), &sclu[bsdia[i]], vloc.data());
    std::memcpy(&diagg[bsdia[i]], vloc.data(), sizeof(double) * b * b);
  }

  int icycle = -1;
  std::vector<int> imask(nsup, -1), lstloc(nsup), iv(nsup, -1), madj(nsup, -1), ibegm(nsup, -1),
      lstupd(nsup), indlst(nsup);
  std::vector<double> fml((size_t)n * blamx, 0.0), fmu((size_t)n * blamx, 0.0);

  const int64_t cap_vals = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)AL.nza()) : -1;
  const int64_t cap_blocks = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)AL.nzja()) : -1;
  std::vector<int> ig(nsup + 1, 0), jg;
  std::vector<int64_t> addrg;
  std::vector<double> gl, gu;
  int ibegfr = 0;
  int64_t nzjrl = 0, nzrtot = 0, nzju = 0, nzjr = 0;
  int64_t compact_mark = std::max<int64_t>(AL.nza(), 1 << 20);

  auto free_r = [&](int irow) {
    int64_t nzjg = ig[ibegfr];
    int64_t nznew = ibegfr != 0 ? addrg[ig[ibegfr]] : 0;
    int64_t beg = ig[ibegfr];
    int ibgfrn = ibegfr;
    nzjrl = 0;
    for (int i = ibegfr; i < irow; ++i) {
      const int bi = bla(i);
      for (int64_t j = beg; j < ig[i + 1]; ++j) {
        int col = jg[j];
        if (col < 0) {
          if (-col < irow) continue;
          ++nzjrl;
          col = -col;
        }
        const int64_t sz = (int64_t)bi * bla(col);
        if (nzjg != j) {
          jg[nzjg] = jg[j];
          std::memmove(&gl[nznew], &gl[addrg[j]], sizeof(double) * sz);
          std::memmove(&gu[nznew], &gu[addrg[j]], sizeof(double) * sz);
        }
        addrg[nzjg] = nznew;
        nznew += sz;
        ++nzjg;
      }
      beg = ig[i + 1];
      ig[i + 1] = (int)nzjg;
      if (nzjrl == 0 && ibgfrn == i) ibgfrn = i + 1;
      iv[i] += (int)(nzjg - beg);
    }
    ibegfr = ibgfrn;
    jg.resize(nzjg);
    addrg.resize(nzjg);
    gl.resize(nznew);
    gu.resize(nznew);
  };

  for (int irow = 0; irow < nsup; ++irow) {
    ++icycle;
    const int bi = bla(irow);
    const int64_t ibsv = AL.sprnd[irow];

    // ---- Ilu2InitRow ----
    int nlist1 = 0;
    lstloc[nlist1++] = irow;
    imask[irow] = icycle;
    std::memcpy(&fml[ibsv * bi], &diagg[bsdia[irow]], sizeof(double) * bi * bi);
    std::memcpy(&fmu[ibsv * bi], &diagg[bsdia[irow]], sizeof(double) * bi * bi);
    for (int j = AL.ia[irow] + 1; j < AL.ia[irow + 1]; ++j) {
      const int jj = AL.ja[j];
      lstloc[nlist1++] = jj;
      imask[jj] = icycle;
      const int bj = bla(jj);
      const int64_t off = (int64_t)AL.sprnd[jj] * bi;
      left_mul(bi, bj, &scll[bsdia[irow]], &AU.a[AU.bsa[j]], aloc.data());
      right_mul_t(bi, bj, aloc.data(), &sclu[bsdia[jj]], vloc.data());
      std::memcpy(&fmu[off], vloc.data(), sizeof(double) * bi * bj);
      left_mul(bi, bj, &sclu[bsdia[irow]], &AL.a[AL.bsa[j]], aloc.data());
      right_mul_t(bi, bj, aloc.data(), &scll[bsdia[jj]], vloc.data());
      std::memcpy(&fml[off], vloc.data(), sizeof(double) * bi * bj);
    }

    // ---- updates by the pending rows ----
    int irwprv = ibegm[irow];
    while (irwprv != -1) {
      const int64_t jpos = iv[irwprv];
      const int icolmn = jg[jpos];
      const int bp = bla(irwprv);
      std::memcpy(lelem.data(), &gl[addrg[jpos]], sizeof(double) * bp * bi);
      std::memcpy(uelem.data(), &gu[addrg[jpos]], sizeof(double) * bp * bi);
      const bool elmisr = icolmn < 0;
      int nupd = 0;
      for (int64_t k = iv[irwprv]; k < ig[irwprv + 1]; ++k) {
        int jc = jg[k];
        if (elmisr && jc < 0) continue;
        if (jc < 0) jc = -jc;
        if (imask[jc] != icycle) {
          imask[jc] = icycle;
          lstloc[nlist1++] = jc;
          const int64_t off = (int64_t)AL.sprnd[jc] * bi;
          std::memset(&fml[off], 0, sizeof(double) * bi * bla(jc));
          std::memset(&fmu[off], 0, sizeof(double) * bi * bla(jc));
        }
        lstupd[nupd] = jc;
        indlst[nupd] = (int)(k - iv[irwprv]);
        ++nupd;
      }
      // U(irow, col) -= Lt(p, irow)^T U(p, col) over all entries, then
      // Lt(irow, col) -= U(p, irow)^T Lt(p, col).
      for (int u = 0; u < nupd; ++u) {
        const int col = lstupd[u];
        sub_tprod(bi, bla(col), bp, lelem.data(), &gu[addrg[iv[irwprv] + indlst[u]]],
                  &fmu[(int64_t)AL.sprnd[col] * bi]);
      }
      for (int u = 0; u < nupd; ++u) {
        const int col = lstupd[u];
        sub_tprod(bi, bla(col), bp, uelem.data(), &gl[addrg[iv[irwprv] + indlst[u]]],
                  &fml[(int64_t)AL.sprnd[col] * bi]);
      }
      irwprv = madj[irwprv];
    }

    // ---- Ich2UpdateTransp ----
    {
      int p = ibegm[irow];
      while (p != -1) {
        const int pnext = madj[p];
        if (iv[p] >= ig[p + 1] - 1) {
          madj[p] = -1;
        } else {
          int col = jg[iv[p] + 1];
          if (col < 0) col = -col;
          madj[p] = ibegm[col];
          ibegm[col] = p;
        }
        iv[p]++;
        p = pnext;
      }
    }

    // ---- Ilu2FiltrRow ----
    {
      int nlist2 = 0;
      double* fdiag = &fmu[ibsv * bi];
      for (int i = 1; i < nlist1; ++i) {
        const int j = lstloc[i];
        const int bj = bla(j);
        const double* fl = &fml[(int64_t)AL.sprnd[j] * bi];
        const double* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
        double amax = 0.0;
        for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(fl[k]));
        for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(fu[k]));
        if (amax >= prm.tau2) {
          lstupd[nlist2++] = j;
        } else {
          double* dj = &diagg[bsdia[j]];
          for (int r = 0; r < bi; ++r)
            for (int c = 0; c < bj; ++c) {
              double v = std::fabs(fl[c * bi + r]);
              double v1 = std::fabs(fu[c * bi + r]);
              if (v1 > v) v = v1;
              v *= prm.theta;
              fdiag[r * bi + r] += v;
              dj[c * bj + c] += v;
            }
        }
      }
      nlist1 = nlist2;
    }

    // ---- Ilu2PivScaleRow ----
    {
      std::memcpy(aloc.data(), &fmu[ibsv * bi], sizeof(double) * bi * bi);
      if (bi > 1) {
        svd_square(bi, aloc.data(), eig.data(), uloc.data(), vtloc.data());
      } else {
        uloc[0] = 1.0;
        if (aloc[0] >= 0.0) { vtloc[0] = 1.0; eig[0] = aloc[0]; }
        else { vtloc[0] = -1.0; eig[0] = -aloc[0]; }
      }
      for (int k = 0; k < bi; ++k) {
        if (eig[k] < prm.pivmin) eig[k] = prm.pivmin;
        eig[k] = std::sqrt(eig[k]);
        eig[k] = 1.0 / eig[k];
      }
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bi; ++c) deleml[c * bi + r] = uloc[r * bi + c] * eig[r];
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bi; ++c) delemu[c * bi + r] = vtloc[c * bi + r] * eig[r];
      for (int i = 0; i < nlist1; ++i) {
        const int j = lstupd[i];
        const int bj = bla(j);
        double* fl = &fml[(int64_t)AL.sprnd[j] * bi];
        double* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
        left_mul(bi, bj, delemu.data(), fl, aloc.data());
        std::memcpy(fl, aloc.data(), sizeof(double) * bi * bj);
        left_mul(bi, bj, deleml.data(), fu, aloc.data());
        std::memcpy(fu, aloc.data(), sizeof(double) * bi * bj);
      }
    }

    int64_t nzloc = (int64_t)bi * bi;
    for (int i = 0; i < nlist1; ++i) nzloc += (int64_t)bi * bla(lstupd[i]);
    int64_t nzjg = (int64_t)jg.size();
    int64_t nznew = (int64_t)gl.size();
    if (cap_vals > 0) {
      if (nzjg + nlist1 + 1 > cap_blocks || nznew + nzloc > cap_vals) {
        free_r(irow);
        nzjg = (int64_t)jg.size();
        nznew = (int64_t)gl.size();
      }
      if (nzjg + nlist1 + 1 > cap_blocks || nznew + nzloc > cap_vals)
        throw Error(kErrWorkspace, " Insufficient workspace in Ilu2");
    } else if (nznew + nzloc > 2 * compact_mark) {
      free_r(irow);
      nzjg = (int64_t)jg.size();
      nznew = (int64_t)gl.size();
      compact_mark = std::max(compact_mark, nznew);
    }

    std::sort(lstupd.begin(), lstupd.begin() + nlist1);
    if (nlist1 != 0) {
      iv[irow] = (int)(nzjg + 1);
      madj[irow] = ibegm[lstupd[0]];
      ibegm[lstupd[0]] = irow;
    }

    // ---- Ilu2StoreRow ----
    jg.push_back(irow);
    addrg.push_back(nznew);
    gl.resize(nznew + nzloc);
    gu.resize(nznew + nzloc);
    for (int r = 0; r < bi; ++r)
      for (int c = 0; c < bi; ++c) {
        gl[nznew + c * bi + r] = deleml[r * bi + c];
        gu[nznew + c * bi + r] = delemu[r * bi + c];
      }
    nznew += (int64_t)bi * bi;
    ++nzju;
    for (int i = 0; i < nlist1; ++i) {
      const int j = lstupd[i];
      const int bj = bla(j);
      const double* fl = &fml[(int64_t)AL.sprnd[j] * bi];
      const double* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
      double amax = 0.0;
      for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(fl[k]));
      for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, std::fabs(fu[k]));
      addrg.push_back(nznew);
      if (amax >= prm.tau1) {
        jg.push_back(j);
        ++nzju;
      } else {
        jg.push_back(-j);
        ++nzjr;
        ++nzjrl;
        nzrtot += (int64_t)bi * bj;
      }
      std::memcpy(&gl[nznew], fl, sizeof(double) * bi * bj);
      std::memcpy(&gu[nznew], fu, sizeof(double) * bi * bj);
      nznew += (int64_t)bi * bj;
    }
    if ((int64_t)jg.size() > INT32_MAX) throw Error(kErrWorkspace, " Insufficient workspace in Ilu2");
    ig[irow + 1] = (int)jg.size();
  }

  // ---- unscale (Ich2StoreScaleU for L and for U) ----
  auto store = [&](const std::vector<double>& g, const std::vector<double>& scl,
                   const std::vector<double>& sclinv, BlockCsr& M) {
    M = BlockCsr();
    M.nsup = nsup;
    M.sprnd = AL.sprnd;
    M.ia.assign(nsup + 1, 0);
    int64_t nz = 0;
    for (int i = 0; i < nsup; ++i)
      for (int64_t j = ig[i]; j < ig[i + 1]; ++j)
        if (jg[j] >= 0) nz += (int64_t)bla(i) * bla(jg[j]);
    M.a.resize(nz);
    nz = 0;
    for (int i = 0; i < nsup; ++i) {
      const int bi = bla(i);
      for (int64_t j = ig[i]; j < ig[i + 1]; ++j) {
        const int jj = jg[j];
        if (jj < 0) continue;
        const int bj = bla(jj);
        if (jj == i) left_mul_t(bi, bi, &scl[bsdia[i]], &g[addrg[j]], &M.a[nz]);
        else right_mul_t(bi, bj, &g[addrg[j]], &sclinv[bsdia[jj]], &M.a[nz]);
        M.ja.push_back(jj);
        M.bsa.push_back(nz);
        nz += (int64_t)bi * bj;
      }
      M.ia[i + 1] = (int)M.ja.size();
    }
    M.set_blamx();
  };
  store(gl, scll, sclinvl, Lout);
  store(gu, sclu, sclinvu, Uout);
  if (st) {
    st->nzu = nzju;
    st->nzr = nzjr;
    st->densu = (double)Lout.nza() / (double)AL.nza();
    st->densr = (double)nzrtot / (double)AL.nza();
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

}  // namespace bcm
