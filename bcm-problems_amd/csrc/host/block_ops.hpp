// Dense block products used by the host factorizations. Blocks are
// column-major. Every output element is a sum over k ascending with each
// product formed from the operands of the reference loops (ich2rs.cpp,
// ilu2rs.cpp), so the rounding sequence is the reference's: the
// implementations vectorize across output rows only, never across k.
#pragma once

namespace bcm {

// out(r,c) = sum_k left(r,k) * mat(k,c); left bl x bl, mat bl x bc.
void left_mul(int bl, int bc, const double* left, const double* mat, double* out);
// out(r,c) = sum_k mat(r,k) * right(c,k); mat br x bc, right bc x bc.
void right_mul_t(int br, int bc, const double* mat, const double* right, double* out);
// out(r,c) = sum_k left(k,r) * mat(k,c); left bl x bl, mat bl x bc.
void left_mul_t(int bl, int bc, const double* left, const double* mat, double* out);
// f(r,c) -= sum_k lt(k,r) * g(k,c), subtracting one product at a time:
// lt is bp x bi (stored with leading dimension bp), g is bp x bj, f is bi x bj.
void sub_tprod(int bi, int bj, int bp, const double* lt, const double* g, double* f);

}  // namespace bcm
