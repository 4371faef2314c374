// Dense block products used by the host factorizations. Blocks are
// column-major; every sum runs over k ascending and every product is formed
// in the operand order of the reference loops (ich2rs.cpp, ilu2rs.cpp), so
// the rounding sequence is the reference's.
#pragma once

namespace bcm {

// out(r,c) = sum_k left(r,k) * mat(k,c); left bl x bl, mat bl x bc.
inline void left_mul(int bl, int bc, const double* left, const double* mat, double* out) {
  for (int r = 0; r < bl; ++r)
    for (int c = 0; c < bc; ++c) {
      double aux = 0.0;
      for (int k = 0; k < bl; ++k) aux += mat[c * bl + k] * left[k * bl + r];
      out[c * bl + r] = aux;
    }
}

// out(r,c) = sum_k mat(r,k) * right(c,k); mat br x bc, right bc x bc.
inline void right_mul_t(int br, int bc, const double* mat, const double* right, double* out) {
  for (int r = 0; r < br; ++r)
    for (int c = 0; c < bc; ++c) {
      double aux = 0.0;
      for (int k = 0; k < bc; ++k) aux += mat[k * br + r] * right[k * bc + c];
      out[c * br + r] = aux;
    }
}

// out(r,c) = sum_k left(k,r) * mat(k,c); left bl x bl, mat bl x bc.
inline void left_mul_t(int bl, int bc, const double* left, const double* mat, double* out) {
  for (int r = 0; r < bl; ++r)
    for (int c = 0; c < bc; ++c) {
      double aux = 0.0;
      for (int k = 0; k < bl; ++k) aux += left[r * bl + k] * mat[c * bl + k];
      out[c * bl + r] = aux;
    }
}

// f(r,c) -= sum_k lt(k,r) * g(k,c), subtracting one product at a time:
// lt is bp x bi (stored with leading dimension bp), g is bp x bj, f is bi x bj.
inline void sub_tprod(int bi, int bj, int bp, const double* lt, const double* g, double* f) {
  for (int r = 0; r < bi; ++r) {
    const double* l = lt + r * bp;
    for (int c = 0; c < bj; ++c) {
      const double* gg = g + c * bp;
      double acc = f[c * bi + r];
      for (int k = 0; k < bp; ++k) acc -= l[k] * gg[k];
      f[c * bi + r] = acc;
    }
  }
}

}  // namespace bcm
