// Second-order incomplete block LU (ILU2) of a complex block matrix on the host.
//
// Restates CSMatrixCS::Ilu2 (ilu2cs.cpp:1213-1567) and its helpers
// Ilu2AllocScale (:26-322, ZGESVD scaling with the sclmin floor),
// Ilu2InitRow (:350-491), Ilu2InitMask (:493-543), Ich2UpdateRow
// (:545-595, applied to U and then to L), Ilu2FiltrRow (:597-694),
// Ilu2PivScaleRow (:696-860, ZGESVD pivots with the pivmin floor),
// Ilu2FreeR (:862-965), Ilu2StoreRow (:967-1085) and Ich2StoreScaleU
// (:1087-1211) for L and for U. The step order is the real ILU2's
// (ilu2.cpp); the differences are the conjugations of the singular vectors
// in the scalings and pivots, and fcttyp: 0 keeps exactly the blocks of the
// pattern of A, otherwise blocks are dropped by the tau1 / tau2 thresholds
// on the largest modulus.
//
// Arithmetic is the reference's dcmplx (complex.h): a product is
// (ac - bd, ad + bc), sums start at zero and add one product at a time,
// moduli are sqrt(re^2 + im^2), and a real diagonal compensation adds to the
// real part only; every block sum runs over k ascending as in the reference
// loops, so the factors equal the reference's bit for bit.
//
// Values: interleaved (re, im) pairs in BlockCsr::a, bsa counting complex
// entries. Input: L stored by columns (the transposed lower blocks, not
// conjugated) and U by rows with identical patterns (combine_lu_complex).
// Output: L by columns and U by rows with inverted diagonal blocks, A ~ L U.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "bcm_host.hpp"

namespace bcm {

namespace {

inline Zc cmul(Zc a, Zc b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline void pleq(Zc& s, Zc p) {
  s.re += p.re;
  s.im += p.im;
}
inline void mneq(Zc& s, Zc p) {
  s.re -= p.re;
  s.im -= p.im;
}
inline double cabs_(Zc z) { return std::sqrt(z.re * z.re + z.im * z.im); }
inline Zc conj_(Zc z) { return {z.re, -z.im}; }
inline Zc cmuld(Zc z, double d) { return {z.re * d, z.im * d}; }
constexpr Zc kCZero{0.0, 0.0};
constexpr Zc kCOne{1.0, 0.0};

// out(r,c) = sum_k left(r,k) * mat(k,c); left bl x bl, mat bl x bc.
void cleft_mul(int bl, int bc, const Zc* left, const Zc* mat, Zc* out) {
  for (int c = 0; c < bc; ++c)
    for (int r = 0; r < bl; ++r) {
      Zc s = kCZero;
      for (int k = 0; k < bl; ++k) pleq(s, cmul(mat[c * bl + k], left[k * bl + r]));
      out[c * bl + r] = s;
    }
}

// out(r,c) = sum_k mat(r,k) * right(c,k); mat br x bc, right bc x bc.
void cright_mul_t(int br, int bc, const Zc* mat, const Zc* right, Zc* out) {
  for (int c = 0; c < bc; ++c)
    for (int r = 0; r < br; ++r) {
      Zc s = kCZero;
      for (int k = 0; k < bc; ++k) pleq(s, cmul(mat[k * br + r], right[k * bc + c]));
      out[c * br + r] = s;
    }
}

// out(r,c) = sum_k left(k,r) * mat(k,c); left bl x bl, mat bl x bc.
void cleft_mul_t(int bl, int bc, const Zc* left, const Zc* mat, Zc* out) {
  for (int c = 0; c < bc; ++c)
    for (int r = 0; r < bl; ++r) {
      Zc s = kCZero;
      for (int k = 0; k < bl; ++k) pleq(s, cmul(left[r * bl + k], mat[c * bl + k]));
      out[c * bl + r] = s;
    }
}

// f(r,c) -= lt(k,r) * g(k,c) for k ascending, one product at a time; lt is
// bp x bi, g bp x bj, f bi x bj.
void csub_tprod(int bi, int bj, int bp, const Zc* lt, const Zc* g, Zc* f) {
  for (int r = 0; r < bi; ++r)
    for (int c = 0; c < bj; ++c) {
      Zc& fr = f[c * bi + r];
      for (int k = 0; k < bp; ++k) mneq(fr, cmul(lt[r * bp + k], g[c * bp + k]));
    }
}

// a = U diag(s) V^H of one square block (ZGESVD, or the reference's own
// 1 x 1 case: s = |a|, U = a / |a|, V = 1).
void block_svd(int b, const Zc* a, double* s, Zc* u, Zc* vt) {
  if (b > 1) {
    svd_square_complex(b, a, s, u, vt);
    return;
  }
  s[0] = cabs_(a[0]);
  if (s[0] > 0.0) {
    u[0] = Zc{a[0].re / s[0], a[0].im / s[0]};
    vt[0] = kCOne;
  } else {
    u[0] = kCOne;
    vt[0] = kCOne;
  }
}

const Zc* zv(const BlockCsr& m) { return reinterpret_cast<const Zc*>(m.a.data()); }

}  // namespace

void ilu2_complex(const BlockCsr& AL, const BlockCsr& AU, const SlvParam& prm, BlockCsr& Lout, BlockCsr& Uout,
                  FactorStats* st) {
  auto t0 = std::chrono::steady_clock::now();
  if (AL.nsup != AU.nsup || AL.ja != AU.ja || AL.ia != AU.ia || AL.sprnd != AU.sprnd || AL.bsa != AU.bsa)
    throw Error(kErrArg, "ilu2: L and U parts must share one pattern");
  const int nsup = AL.nsup;
  const int n = AL.n();
  auto bla = [&](int i) { return AL.sprnd[i + 1] - AL.sprnd[i]; };
  int64_t nza = 0;
  for (int i = 0; i < nsup; ++i) {
    if (AL.ia[i] >= AL.ia[i + 1] || AL.ja[AL.ia[i]] != i)
      throw Error(kErrArg, "ilu2: every row must start with its diagonal block");
    for (int j = AL.ia[i]; j < AL.ia[i + 1]; ++j) {
      if (AL.bsa[j] != nza) throw Error(kErrArg, "ilu2: blocks must be stored in ja order");
      nza += (int64_t)bla(i) * bla(AL.ja[j]);
    }
  }
  if ((int64_t)AL.a.size() != 2 * nza || (int64_t)AU.a.size() != 2 * nza)
    throw Error(kErrArg, "ilu2: complex values must hold two doubles per entry");
  const Zc* al = zv(AL);
  const Zc* au = zv(AU);
  int blamx = 1;
  for (int i = 0; i < nsup; ++i) blamx = std::max(blamx, bla(i));

  std::vector<int64_t> bsdia(nsup + 1);
  int64_t nzd = 0;
  for (int i = 0; i < nsup; ++i) {
    bsdia[i] = nzd;
    nzd += (int64_t)bla(i) * bla(i);
  }
  bsdia[nsup] = nzd;
  std::vector<Zc> scll(nzd), sclu(nzd), sclinvl(nzd), sclinvu(nzd), diagg(nzd);
  const int bb = blamx * blamx;
  std::vector<Zc> aloc(bb), vloc(bb), uloc(bb), vtloc(bb), deleml(bb), delemu(bb), lelem(bb), uelem(bb);
  std::vector<double> eig(blamx);

  // ---- scaling (Ilu2AllocScale) ----
  for (int i = 0; i < nsup; ++i) {
    const int b = bla(i);
    const Zc* aii = au + AU.bsa[AU.ia[i]];
    block_svd(b, aii, eig.data(), uloc.data(), vtloc.data());
    for (int k = 0; k < b; ++k)
      if (eig[k] < prm.sclmin) eig[k] = prm.sclmin;
    for (int k = 0; k < b; ++k) eig[k] = std::sqrt(eig[k]);
    Zc* sil = &sclinvl[bsdia[i]];
    Zc* siu = &sclinvu[bsdia[i]];
    Zc* sl = &scll[bsdia[i]];
    Zc* su = &sclu[bsdia[i]];
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) sil[c * b + r] = cmuld(uloc[c * b + r], eig[c]);
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) siu[c * b + r] = cmuld(vtloc[r * b + c], eig[c]);
    for (int k = 0; k < b; ++k) eig[k] = 1.0 / eig[k];
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) sl[c * b + r] = cmuld(conj_(uloc[r * b + c]), eig[r]);
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) su[c * b + r] = cmuld(conj_(vtloc[c * b + r]), eig[r]);
    cleft_mul(b, b, sl, aii, aloc.data());
    cright_mul_t(b, b, aloc.data(), su, vloc.data());
    std::copy(vloc.begin(), vloc.begin() + (size_t)b * b, diagg.begin() + bsdia[i]);
  }

  int icycle = -1, icycleini = -1;
  std::vector<int> imask(nsup, -1), imaskini(nsup, -1), lstloc(nsup), iv(nsup, -1), madj(nsup, -1),
      ibegm(nsup, -1), lstupd(nsup), indlst(nsup);
  std::vector<Zc> fml((size_t)n * blamx, kCZero), fmu((size_t)n * blamx, kCZero);
  const bool pattern_only = prm.fcttyp == 0;
  auto mark_pattern = [&](int irow) {  // imaskini over the columns of row irow of A
    ++icycleini;
    for (int j = AL.ia[irow]; j < AL.ia[irow + 1]; ++j) imaskini[AL.ja[j]] = icycleini;
  };

  const int64_t cap_vals = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)nza) : -1;
  const int64_t cap_blocks = prm.memory > 0.0 ? (int64_t)(prm.memory * (double)AL.nzja()) : -1;
  std::vector<int> ig(nsup + 1, 0), jg;
  std::vector<int64_t> addrg;
  std::vector<Zc> gl, gu;
  int ibegfr = 0;
  int64_t nzjrl = 0, nzrtot = 0, nzju = 0, nzjr = 0;
  int64_t compact_mark = std::max<int64_t>(nza, 1 << 20);

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
          std::memmove(&gl[nznew], &gl[addrg[j]], sizeof(Zc) * sz);
          std::memmove(&gu[nznew], &gu[addrg[j]], sizeof(Zc) * sz);
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
    std::copy(diagg.begin() + bsdia[irow], diagg.begin() + bsdia[irow] + (size_t)bi * bi, fml.begin() + ibsv * bi);
    std::copy(diagg.begin() + bsdia[irow], diagg.begin() + bsdia[irow] + (size_t)bi * bi, fmu.begin() + ibsv * bi);
    for (int j = AL.ia[irow] + 1; j < AL.ia[irow + 1]; ++j) {
      const int jj = AL.ja[j];
      lstloc[nlist1++] = jj;
      imask[jj] = icycle;
      const int bj = bla(jj);
      const int64_t off = (int64_t)AL.sprnd[jj] * bi;
      cleft_mul(bi, bj, &scll[bsdia[irow]], au + AU.bsa[j], aloc.data());
      cright_mul_t(bi, bj, aloc.data(), &sclu[bsdia[jj]], vloc.data());
      std::copy(vloc.begin(), vloc.begin() + (size_t)bi * bj, fmu.begin() + off);
      cleft_mul(bi, bj, &sclu[bsdia[irow]], al + AL.bsa[j], aloc.data());
      cright_mul_t(bi, bj, aloc.data(), &scll[bsdia[jj]], vloc.data());
      std::copy(vloc.begin(), vloc.begin() + (size_t)bi * bj, fml.begin() + off);
    }

    // ---- updates by the pending rows (Ilu2InitMask, Ich2UpdateRow for U then L) ----
    int irwprv = ibegm[irow];
    while (irwprv != -1) {
      const int64_t jpos = iv[irwprv];
      const int icolmn = jg[jpos];
      const int bp = bla(irwprv);
      std::copy(gl.begin() + addrg[jpos], gl.begin() + addrg[jpos] + (size_t)bp * bi, lelem.begin());
      std::copy(gu.begin() + addrg[jpos], gu.begin() + addrg[jpos] + (size_t)bp * bi, uelem.begin());
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
          std::fill(fml.begin() + off, fml.begin() + off + (size_t)bi * bla(jc), kCZero);
          std::fill(fmu.begin() + off, fmu.begin() + off + (size_t)bi * bla(jc), kCZero);
        }
        lstupd[nupd] = jc;
        indlst[nupd] = (int)(k - iv[irwprv]);
        ++nupd;
      }
      for (int u = 0; u < nupd; ++u) {
        const int col = lstupd[u];
        csub_tprod(bi, bla(col), bp, lelem.data(), &gu[addrg[iv[irwprv] + indlst[u]]],
                   &fmu[(int64_t)AL.sprnd[col] * bi]);
      }
      for (int u = 0; u < nupd; ++u) {
        const int col = lstupd[u];
        csub_tprod(bi, bla(col), bp, uelem.data(), &gl[addrg[iv[irwprv] + indlst[u]]],
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
      if (pattern_only) mark_pattern(irow);
      int nlist2 = 0;
      Zc* fdiag = &fmu[ibsv * bi];
      for (int i = 1; i < nlist1; ++i) {
        const int j = lstloc[i];
        const int bj = bla(j);
        const Zc* fl = &fml[(int64_t)AL.sprnd[j] * bi];
        const Zc* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
        double amax = 0.0;
        for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, cabs_(fl[k]));
        for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, cabs_(fu[k]));
        const bool large = pattern_only ? imaskini[j] == icycleini : amax >= prm.tau2;
        if (large) {
          lstupd[nlist2++] = j;
        } else {
          Zc* dj = &diagg[bsdia[j]];
          for (int r = 0; r < bi; ++r)
            for (int c = 0; c < bj; ++c) {
              double v = cabs_(fl[c * bi + r]);
              const double v1 = cabs_(fu[c * bi + r]);
              if (v1 > v) v = v1;
              v *= prm.theta;
              fdiag[r * bi + r].re += v;
              dj[c * bj + c].re += v;
            }
        }
      }
      nlist1 = nlist2;
    }

    // ---- Ilu2PivScaleRow ----
    {
      std::copy(fmu.begin() + ibsv * bi, fmu.begin() + ibsv * bi + (size_t)bi * bi, aloc.begin());
      block_svd(bi, aloc.data(), eig.data(), uloc.data(), vtloc.data());
      for (int k = 0; k < bi; ++k) {
        if (eig[k] < prm.pivmin) eig[k] = prm.pivmin;
        eig[k] = std::sqrt(eig[k]);
        eig[k] = 1.0 / eig[k];
      }
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bi; ++c) deleml[c * bi + r] = cmuld(conj_(uloc[r * bi + c]), eig[r]);
      for (int r = 0; r < bi; ++r)
        for (int c = 0; c < bi; ++c) delemu[c * bi + r] = cmuld(conj_(vtloc[c * bi + r]), eig[r]);
      for (int i = 0; i < nlist1; ++i) {
        const int j = lstupd[i];
        const int bj = bla(j);
        Zc* fl = &fml[(int64_t)AL.sprnd[j] * bi];
        Zc* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
        cleft_mul(bi, bj, delemu.data(), fl, aloc.data());
        std::copy(aloc.begin(), aloc.begin() + (size_t)bi * bj, fl);
        cleft_mul(bi, bj, deleml.data(), fu, aloc.data());
        std::copy(aloc.begin(), aloc.begin() + (size_t)bi * bj, fu);
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
    if (pattern_only) mark_pattern(irow);
    for (int i = 0; i < nlist1; ++i) {
      const int j = lstupd[i];
      const int bj = bla(j);
      const Zc* fl = &fml[(int64_t)AL.sprnd[j] * bi];
      const Zc* fu = &fmu[(int64_t)AL.sprnd[j] * bi];
      double amax = 0.0;
      for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, cabs_(fl[k]));
      for (int k = 0; k < bi * bj; ++k) amax = std::max(amax, cabs_(fu[k]));
      addrg.push_back(nznew);
      const bool first = pattern_only ? imaskini[j] == icycleini : amax >= prm.tau1;
      if (first) {
        jg.push_back(j);
        ++nzju;
      } else {
        jg.push_back(-j);
        ++nzjr;
        ++nzjrl;
        nzrtot += (int64_t)bi * bj;
      }
      std::copy(fl, fl + (size_t)bi * bj, gl.begin() + nznew);
      std::copy(fu, fu + (size_t)bi * bj, gu.begin() + nznew);
      nznew += (int64_t)bi * bj;
    }
    if ((int64_t)jg.size() > INT32_MAX) throw Error(kErrWorkspace, " Insufficient workspace in Ilu2");
    ig[irow + 1] = (int)jg.size();
  }

  // ---- unscale (Ich2StoreScaleU for L and for U) ----
  auto store = [&](const std::vector<Zc>& g, const std::vector<Zc>& scl, const std::vector<Zc>& sclinv,
                   BlockCsr& M) {
    M = BlockCsr();
    M.nsup = nsup;
    M.sprnd = AL.sprnd;
    M.ia.assign(nsup + 1, 0);
    int64_t nz = 0;
    for (int i = 0; i < nsup; ++i)
      for (int64_t j = ig[i]; j < ig[i + 1]; ++j)
        if (jg[j] >= 0) nz += (int64_t)bla(i) * bla(jg[j]);
    M.a.resize(2 * nz);
    Zc* out = reinterpret_cast<Zc*>(M.a.data());
    nz = 0;
    for (int i = 0; i < nsup; ++i) {
      const int bi = bla(i);
      for (int64_t j = ig[i]; j < ig[i + 1]; ++j) {
        const int jj = jg[j];
        if (jj < 0) continue;
        const int bj = bla(jj);
        if (jj == i) cleft_mul_t(bi, bi, &scl[bsdia[i]], &g[addrg[j]], out + nz);
        else cright_mul_t(bi, bj, &g[addrg[j]], &sclinv[bsdia[jj]], out + nz);
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
    st->densu = (double)(Lout.a.size() / 2) / (double)nza;
    st->densr = (double)nzrtot / (double)nza;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

}  // namespace bcm
