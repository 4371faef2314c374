// Dense block products of the host factorizations (declared in
// block_ops.hpp). The loops run k outermost per output column and rows
// innermost, so each output element sees exactly the scalar sequence
//   aux = 0; for k ascending: aux += a_k * b_k      (or f -= a_k * b_k)
// of the reference's loops while the compiler vectorizes across rows.
// Compiled with -ffp-contract=off: no fused multiply-add changes a rounding.
// Row counts 9 and 24 (the block sizes of the synthetic systems) keep the
// column accumulator in registers; other sizes use the same loops on a
// stack column.
#include "block_ops.hpp"

#include <cstddef>
#include <vector>

namespace bcm {

namespace {

constexpr int kMaxStack = 64;

// acc[r] (+|-)= sum_k colsrc(k)[r] * scal(k) for the columns of one output.
template <int BI>
inline void col_accumulate(int nk, const double* rows, int ldrows, const double* scal, double* out, bool subtract) {
  double fv[BI];
  for (int r = 0; r < BI; ++r) fv[r] = out[r];
  if (subtract) {
    for (int k = 0; k < nk; ++k) {
      const double s = scal[k];
      const double* lr = rows + (long)k * ldrows;
      for (int r = 0; r < BI; ++r) fv[r] -= lr[r] * s;
    }
  } else {
    for (int k = 0; k < nk; ++k) {
      const double s = scal[k];
      const double* lr = rows + (long)k * ldrows;
      for (int r = 0; r < BI; ++r) fv[r] += lr[r] * s;
    }
  }
  for (int r = 0; r < BI; ++r) out[r] = fv[r];
}

inline void col_accumulate_n(int bi, int nk, const double* rows, int ldrows, const double* scal, double* out,
                             bool subtract) {
  if (subtract) {
    for (int k = 0; k < nk; ++k) {
      const double s = scal[k];
      const double* lr = rows + (long)k * ldrows;
      for (int r = 0; r < bi; ++r) out[r] -= lr[r] * s;
    }
  } else {
    for (int k = 0; k < nk; ++k) {
      const double s = scal[k];
      const double* lr = rows + (long)k * ldrows;
      for (int r = 0; r < bi; ++r) out[r] += lr[r] * s;
    }
  }
}

// out column c (bi values) (+|-)= rows(bi x nk, leading dim ldrows) * scal(nk), k ascending.
inline void col_op(int bi, int nk, const double* rows, int ldrows, const double* scal, double* out, bool subtract) {
  switch (bi) {
    case 9: col_accumulate<9>(nk, rows, ldrows, scal, out, subtract); break;
    case 24: col_accumulate<24>(nk, rows, ldrows, scal, out, subtract); break;
    default: col_accumulate_n(bi, nk, rows, ldrows, scal, out, subtract); break;
  }
}

// dst(k, r) = src(r, k), with the source
// element (r, k) at src[r * lds + k] (r < nr, k < nk) and dst[k * nr + r].
inline void transpose(int nr, int nk, const double* src, int lds, double* dst) {
  for (int r = 0; r < nr; ++r)
    for (int k = 0; k < nk; ++k) dst[(long)k * nr + r] = src[(long)r * lds + k];
}

}  // namespace

__attribute__((target_clones("avx2", "default")))
void left_mul(int bl, int bc, const double* left, const double* mat, double* out) {
  // out(:,c) = 0 + sum_k left(:,k) * mat(k,c); left(r,k) = left[k*bl + r]
  for (int c = 0; c < bc; ++c) {
    double* oc = out + (long)c * bl;
    for (int r = 0; r < bl; ++r) oc[r] = 0.0;
    col_op(bl, bl, left, bl, mat + (long)c * bl, oc, false);
  }
}

__attribute__((target_clones("avx2", "default")))
void right_mul_t(int br, int bc, const double* mat, const double* right, double* out) {
  // out(:,c) = 0 + sum_k mat(:,k) * right(c,k); mat(r,k) = mat[k*br + r], right(c,k) = right[k*bc + c]
  double sbuf[kMaxStack];
  std::vector<double> heap;
  double* sc = sbuf;
  if (bc > kMaxStack) { heap.resize(bc); sc = heap.data(); }
  for (int c = 0; c < bc; ++c) {
    for (int k = 0; k < bc; ++k) sc[k] = right[(long)k * bc + c];
    double* oc = out + (long)c * br;
    for (int r = 0; r < br; ++r) oc[r] = 0.0;
    col_op(br, bc, mat, br, sc, oc, false);
  }
}

__attribute__((target_clones("avx2", "default")))
void left_mul_t(int bl, int bc, const double* left, const double* mat, double* out) {
  // out(:,c) = 0 + sum_k left(k,:) * mat(k,c); left(k,r) = left[r*bl + k]
  std::vector<double> lt((std::size_t)bl * bl);
  transpose(bl, bl, left, bl, lt.data());
  for (int c = 0; c < bc; ++c) {
    double* oc = out + (long)c * bl;
    for (int r = 0; r < bl; ++r) oc[r] = 0.0;
    col_op(bl, bl, lt.data(), bl, mat + (long)c * bl, oc, false);
  }
}

__attribute__((target_clones("avx2", "default")))
void sub_tprod(int bi, int bj, int bp, const double* lt, const double* g, double* f) {
  // f(:,c) -= sum_k lt(k,:) * g(k,c); lt(k,r) = lt[r*bp + k]
  double sbuf[kMaxStack * kMaxStack];
  std::vector<double> heap;
  double* t = sbuf;
  if ((long)bi * bp > kMaxStack * kMaxStack) { heap.resize((std::size_t)bi * bp); t = heap.data(); }
  transpose(bi, bp, lt, bp, t);
  for (int c = 0; c < bj; ++c) col_op(bi, bp, t, bi, g + (long)c * bp, f + (long)c * bi, true);
}

}  // namespace bcm
