// Complex block matrices and ILU2 factors of the complex solve phase
// (CSMatrixCS, smatrix.h:1499-1700; MvmA / SolveL / SolveU, mvmc.cpp:160,
// 1723, 2506) held as their real equivalents, so that every real kernel and
// Krylov driver of this library applies the complex operator:
//
//   a complex block C (p x q) becomes the real 2p x 2q block
//       [ Re C  -Im C ]
//       [ Im C   Re C ]
//   and a complex vector block z (p values) the real block [ Re z ; Im z ],
//
// so that the real product of the two is the real equivalent of C z. The
// inverse of a real-equivalent block is the real equivalent of the inverse,
// so the pre-inverted diagonal blocks of the factors carry over.
//
// Storage follows the reference's: the complex values are interleaved
// (re, im) per entry (dcmplx), blocks column-major in ja order. The blocks
// of L, stored by columns, hold the transpose of the logical block (SolveL
// reads entry (k, j) of block L_ki at kk * b_i + j), so their real
// equivalent is RE(M)^T = [[Re C, Im C], [-Im C, Re C]] with C = M^T, the
// block as stored; U (by rows) and A store blocks as they are.
//
// Complex arithmetic accumulates (Re, Im) pairs where the real form sums the
// real and imaginary products in separate column runs, so products agree
// with the reference's complex ones to rounding, not bit for bit.
#include "bcm_host.hpp"

namespace bcm {

namespace {

// Real equivalent of one block: src has p x q complex entries column-major
// (2 doubles each), dst gets 2p x 2q doubles column-major.
void expand_block(int p, int q, const double* src, bool transposed, double* dst) {
  const int P = 2 * p;
  const double s = transposed ? -1.0 : 1.0;  // sign of the lower-left imaginary part
  for (int c = 0; c < q; ++c)
    for (int r = 0; r < p; ++r) {
      const double re = src[2 * ((size_t)c * p + r)], im = src[2 * ((size_t)c * p + r) + 1];
      dst[(size_t)c * P + r] = re;
      dst[(size_t)c * P + r + p] = s * im;
      dst[(size_t)(c + q) * P + r] = -s * im;
      dst[(size_t)(c + q) * P + r + p] = re;
    }
}

}  // namespace

BlockCsr real_equivalent(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                         bool transposed_blocks) {
  if (nsup < 0) throw Error(kErrArg, "complex matrix: negative size");
  BlockCsr r;
  r.nsup = nsup;
  r.sprnd.assign(nsup + 1, 0);
  for (int i = 0; i < nsup; ++i) {
    const int b = sprnd[i + 1] - sprnd[i];
    if (b <= 0) throw Error(kErrArg, "complex matrix: empty block row");
    r.sprnd[i + 1] = r.sprnd[i] + 2 * b;
  }
  r.ia.assign(ia, ia + nsup + 1);
  if (r.ia[0] != 0) throw Error(kErrArg, "complex matrix: ia[0] must be 0");
  const int nzja = r.ia[nsup];
  r.ja.assign(ja, ja + nzja);
  r.bsa.resize(nzja);
  int64_t off = 0, coff = 0;  // real and complex value offsets of the block
  for (int i = 0; i < nsup; ++i)
    for (int j = r.ia[i]; j < r.ia[i + 1]; ++j) {
      if (r.ja[j] < 0 || r.ja[j] >= nsup) throw Error(kErrArg, "complex matrix: block column out of range");
      r.bsa[j] = off;
      const int p = sprnd[i + 1] - sprnd[i], q = sprnd[r.ja[j] + 1] - sprnd[r.ja[j]];
      off += (int64_t)4 * p * q;
      coff += (int64_t)p * q;
    }
  r.a.resize((size_t)off);
  coff = 0;
  for (int i = 0; i < nsup; ++i)
    for (int j = r.ia[i]; j < r.ia[i + 1]; ++j) {
      const int p = sprnd[i + 1] - sprnd[i], q = sprnd[r.ja[j] + 1] - sprnd[r.ja[j]];
      expand_block(p, q, a + 2 * coff, transposed_blocks, r.a.data() + r.bsa[j]);
      coff += (int64_t)p * q;
    }
  r.set_blamx();
  return r;
}

void split_complex(const BlockCsr& c, BlockCsr& re, BlockCsr& im) {
  if (c.a.size() % 2) throw Error(kErrArg, "complex matrix: odd number of values");
  re = c;
  re.a.resize(c.a.size() / 2);
  im = re;
  for (size_t k = 0; k < re.a.size(); ++k) {
    re.a[k] = c.a[2 * k];
    im.a[k] = c.a[2 * k + 1];
  }
}

BlockCsr merge_complex(const BlockCsr& re, const BlockCsr& im) {
  if (re.ja != im.ja || re.ia != im.ia || re.bsa != im.bsa || re.a.size() != im.a.size())
    throw Error(kErrArg, "complex matrix: real and imaginary parts differ in pattern");
  BlockCsr c = re;
  c.a.resize(2 * re.a.size());
  for (size_t k = 0; k < re.a.size(); ++k) {
    c.a[2 * k] = re.a[k];
    c.a[2 * k + 1] = im.a[k];
  }
  return c;
}

void combine_lu_complex(const BlockCsr& a, BlockCsr& l, BlockCsr& u) {
  BlockCsr re, im, lre, lim, ure, uim;
  split_complex(a, re, im);
  combine_lu(re, lre, ure);
  combine_lu(im, lim, uim);
  l = merge_complex(lre, lim);
  u = merge_complex(ure, uim);
}

void complex_to_real(int nsup, const int* sprnd, const double* z, double* x, int nrhs) {
  const int64_t n = sprnd[nsup];
  for (int q = 0; q < nrhs; ++q)
    for (int i = 0; i < nsup; ++i) {
      const int b = sprnd[i + 1] - sprnd[i];
      const double* zi = z + 2 * ((size_t)q * n + sprnd[i]);
      double* xi = x + (size_t)q * 2 * n + 2 * (size_t)sprnd[i];
      for (int k = 0; k < b; ++k) {
        xi[k] = zi[2 * k];
        xi[b + k] = zi[2 * k + 1];
      }
    }
}

void real_to_complex(int nsup, const int* sprnd, const double* x, double* z, int nrhs) {
  const int64_t n = sprnd[nsup];
  for (int q = 0; q < nrhs; ++q)
    for (int i = 0; i < nsup; ++i) {
      const int b = sprnd[i + 1] - sprnd[i];
      double* zi = z + 2 * ((size_t)q * n + sprnd[i]);
      const double* xi = x + (size_t)q * 2 * n + 2 * (size_t)sprnd[i];
      for (int k = 0; k < b; ++k) {
        zi[2 * k] = xi[k];
        zi[2 * k + 1] = xi[b + k];
      }
    }
}

}  // namespace bcm
