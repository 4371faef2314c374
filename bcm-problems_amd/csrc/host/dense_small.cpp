// Small dense eigen/singular-value decompositions used by the IC2/ILU2 host
// factorizations for the scaling of every diagonal block and for the block
// pivots. The reference calls LAPACK 3.0 DSYEV / DGESVD through
// lapackd.cpp:865-905 with a workspace of 10 n doubles (ich2rs.cpp:56,
// ilu2rs.cpp:83); these wrappers call the restatements in lapack30.cpp the
// same way, and turn a nonzero INFO into the reference's error
// (ich2rs.cpp:112, ilu2rs.cpp:131). The complex ILU2 calls ZGESVD directly
// with 10 n complex values and 10 n doubles of workspace (ilu2cs.cpp:71-83,
// 132-136).
#include "bcm_host.hpp"

#include <cstring>
#include <vector>

namespace bcm {

void sym_eig(int n, double* a, double* w) {
  std::vector<double> work((size_t)10 * n + 1);
  const int info = lapack30::dsyev_vu(n, a, n, w, work.data(), 10 * n);
  if (info != 0) throw Error(kErrInternal, "Error in the Lapack routine DSYEV");
}

void svd_square(int n, const double* a, double* s, double* u, double* vt) {
  std::vector<double> acopy(a, a + (size_t)n * n), work((size_t)10 * n + 1);
  const int info = lapack30::dgesvd_aa(n, acopy.data(), n, s, u, n, vt, n, work.data(), 10 * n);
  if (info != 0) throw Error(kErrInternal, "Error in the Lapack routine DGESVD");
}

void svd_square_complex(int n, const Zc* a, double* s, Zc* u, Zc* vt) {
  std::vector<Zc> acopy(a, a + (size_t)n * n), work((size_t)10 * n + 1);
  std::vector<double> rwork((size_t)10 * n + 1);
  const int info = lapack30::zgesvd_aa(n, acopy.data(), n, s, u, n, vt, n, work.data(), 10 * n, rwork.data());
  if (info != 0) throw Error(kErrInternal, "Error in the Lapack routine ZGESVD");
}

}  // namespace bcm
