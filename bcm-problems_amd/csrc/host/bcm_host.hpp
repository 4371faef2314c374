// Host-side data model of libbcmhip: the block-CSR container, the solver
// parameters and the host stages that precede the GPU solve (assembly,
// ordering, L/U split and the IC2/ILU2 factorizations).
//
// Reference counterparts (Bslib):
//   BlockCsr   <- CSMatrixRS (smatrix.h:1040-1048): nsupr, sprndr/sprndc, ia, ja,
//                 bsa, a; blocks column-major: element (r,c) of block j at
//                 a[bsa[j] + c*bla_i + r] (mvm.cpp:2560).
//   SlvParam   <- CSlvParam (slvparam.h:19-214), solve-phase subset.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace bcm {

// Status codes shared with the C ABI (include/bcmhip.h).
enum Status : int {
  kOk = 0,
  kErrArg = 1,          // invalid argument / shape
  kErrNotSpd = 2,       // negative eigenvalue in IC2 scaling or pivot
  kErrWorkspace = 3,    // factor would exceed the workspace bound
  kErrBreakdown = 4,    // Krylov breakdown
  kErrHip = 5,          // HIP runtime failure
  kErrNoConv = 6,       // iteration cap reached without convergence
  kErrNoDevice = 7,     // no GPU / library built without device support
  kErrInternal = 8,
};

// Exception carrying a status code and the reference-style message
// (the reference throws C strings such as " Negative pivot in ICH2 ").
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct BlockCsr {
  int nsup = 0;                // number of block rows (== block columns)
  int blamx = 0;               // largest block size
  std::vector<int> sprnd;      // nsup+1 prefix sums of block sizes
  std::vector<int> ia;         // nsup+1 row pointers into ja
  std::vector<int> ja;         // block column indices, ascending per row
  std::vector<int64_t> bsa;    // value offset of every block
  std::vector<double> a;       // block values, column-major per block

  int n() const { return nsup ? sprnd[nsup] : 0; }
  int nzja() const { return nsup ? ia[nsup] : 0; }
  int64_t nza() const { return (int64_t)a.size(); }
  int bla(int i) const { return sprnd[i + 1] - sprnd[i]; }
  void set_blamx();
  // Checks the structural invariants (sorted ja, consistent bsa); throws on violation.
  void validate(bool require_diag_first) const;
};

// One complex value as the reference stores it (dcmplx, complex.h:17-19):
// real part, then imaginary part; arrays of them are interleaved doubles.
struct Zc {
  double re, im;
};

struct SlvParam {
  int msglev = 0;
  int ordtype = 1;       // 1: RCM profile ordering, 0: identity
  int fcttyp = 1;        // complex ILU2 only: 0 keeps the pattern of A, else thresholds (slvparam.h:33)
  double memory = -1.0;  // workspace bound in units of nza (<=0: unbounded)
  double tau1 = 1.0e-2;  // first-order threshold (stored in the factor)
  double tau2 = 1.0e-4;  // second-order threshold (kept for updates only)
  double theta = 1.0e-1; // diagonal compensation of dropped blocks
  double sclmin = 1.0e-8;
  double pivmin = 1.0e-8;
  int ittype = 2;
  int sttype = 2;
  int niter = 1000;
  int ichk = 5;
  double eps = 1.0e-9;
  // outputs
  int nitperf = 0;
  double resfin = 0.0;
  double timefct = 0.0;
  double timeitr = 0.0;
  double density = 0.0;
  double density2 = 0.0;
};

// ---- ordering and structural transforms (order.cpp) ----

// Symmetrized block pattern (union of A and A^T rows, ascending, diagonal kept).
void symm_pattern(const BlockCsr& a, std::vector<int>& ias, std::vector<int>& jas);
// SPARSPAK reverse Cuthill-McKee on a 0-based adjacency; returns perm[new] = old.
void genrcm(int n, const std::vector<int>& xadj, const std::vector<int>& adj, std::vector<int>& perm);
// Profile ordering as in OrderPrfMtr(1): order[old] = new.
std::vector<int> rcm_order(const BlockCsr& a);

// ---- complex matrices as real equivalents (complex.cpp) ----

// Real equivalent of a complex block matrix (interleaved re/im values, blocks
// column-major in ja order): block rows of 2 b_i, each complex block C as
// [[Re C, -Im C], [Im C, Re C]], or, for blocks stored transposed (L of an
// ILU2 factor, by columns), [[Re C, Im C], [-Im C, Re C]].
BlockCsr real_equivalent(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                         bool transposed_blocks);
// Complex vectors (interleaved, n x nrhs) to and from the real-equivalent
// layout (per block row [Re z; Im z], 2n x nrhs).
void complex_to_real(int nsup, const int* sprnd, const double* z, double* x, int nrhs);
void real_to_complex(int nsup, const int* sprnd, const double* x, double* z, int nrhs);

// ---- nested dissection and the parallel solver's tree (nd.cpp) ----

// Nested-dissection ordering of the symmetrized block pattern with the layout
// METIS_NodeND gives OrderPrfMtr(-1) (ordmatr.cpp:28-85): each part ordered
// recursively, then its separator. order[old] = new.
std::vector<int> nd_order(const BlockCsr& a);
// CSMatrix::FindSeparator (smatrix.cpp:4355-4488) on a symmetric pattern of n
// rows: ind1 = first row of part 2, ind2 = first separator row; -1 if none.
void find_separator(int n, const int* ia, const int* ja, int& ind1, int& ind2);
// Blocks of the binary dissection tree over nproc (a power of two) virtual
// ranks, as CSMatrix::PartBinaryTreeND (ordmatr.cpp:1204-1396, 1698-1805) and
// CTree::PartitionTree (tree.cpp:2414-2455) derive them from an ND-ordered
// matrix: blocks in matrix order (leaves and separators), each with the rank
// owning it (a separator: the first rank of its subtree).
struct NdTree {
  int nproc = 1;
  std::vector<int> blks, blk2cpu;
  std::vector<std::pair<int, int>> leaf_rows, separators;  // [first, last + 1) block rows
};
NdTree part_binary_tree_nd(const BlockCsr& ordered, int nproc);
// Permuted copy, order[old] = new, columns re-sorted ascending per row.
BlockCsr ord_mtr(const BlockCsr& a, const std::vector<int>& order);
// Full symmetric matrix from upper-triangular storage (lower blocks transposed).
BlockCsr symm_mtr(const BlockCsr& upper);
// Symmetric permutation of an upper-stored matrix, returned upper-stored.
BlockCsr ord_mtr_symm(const BlockCsr& upper, const std::vector<int>& order);
// Block transpose (pattern and block contents).
BlockCsr transp_mtr(const BlockCsr& a);
// Extend pattern of `a` with zero blocks to the union with `b` (same nsup).
BlockCsr extend_sparsity(const BlockCsr& a, const BlockCsr& b);
// Split into L stored by columns (= transposed lower part) and U by rows, both
// padded to the union pattern.
void combine_lu(const BlockCsr& a, BlockCsr& l, BlockCsr& u);
// Block-permute a vector with nrhs columns of length n: 'D' applies order, 'I' inverts it.
void ord_vect(char type, const std::vector<int>& sprnd_old, const std::vector<int>& sprnd_new,
              const std::vector<int>& order, double* v, int n, int nrhs);
// Upper-triangular part (j >= i) of a full matrix.
BlockCsr upper_part(const BlockCsr& a);

// ---- binary files in the reference's format (binio.cpp) ----

// CSMatrixRS::WriteMtrBin (smatrixrs.cpp:3210-3227) layout.
void write_mtr_bin(const BlockCsr& a, const std::string& path);
// Reads that layout back (square block matrices, identity row list, blocks in ja order).
BlockCsr read_mtr_bin(const std::string& path);
// CSVector::WriteVectBin (svector.cpp:537-546) of an n x nrhs vector (one part).
void write_vect_bin(const std::string& path, int n, int nrhs, const double* v);
// CSVector::ReadVectBin (svector.cpp:488-507): sizes, and the values when v != null.
void read_vect_bin(const std::string& path, int* n, int* nrhs, std::vector<double>* v);

// ---- small dense kernels (lapack30.cpp, dense_small.cpp) ----

namespace lapack30 {
// DSYEV with JOBZ = 'V', UPLO = 'U' (LAPACK 3.0): eigenvalues ascending in w,
// eigenvectors in the columns of a. work holds lwork >= 3n - 1 doubles.
// Returns INFO (0, > 0: DSTEQR did not converge, < 0: bad argument).
int dsyev_vu(int n, double* a, int lda, double* w, double* work, int lwork);
// DGESVD with JOBU = JOBVT = 'A' on a square n x n matrix (LAPACK 3.0):
// a = U diag(s) VT, s descending; a is overwritten. work holds lwork >= 5n
// doubles. Returns INFO (> 0: DBDSQR did not converge, < 0: bad argument).
int dgesvd_aa(int n, double* a, int lda, double* s, double* u, int ldu, double* vt, int ldvt, double* work,
              int lwork);
// ZGESVD with JOBU = JOBVT = 'A' on a square n x n complex matrix (LAPACK
// 3.0): a = U diag(s) VT with VT = V^H, s descending; a is overwritten. work
// holds lwork >= 3n complex values, rwork 5n doubles. Returns INFO.
int zgesvd_aa(int n, Zc* a, int lda, double* s, Zc* u, int ldu, Zc* vt, int ldvt, Zc* work, int lwork,
              double* rwork);
}  // namespace lapack30

// Symmetric eigen-decomposition of an n x n column-major matrix: eigenvalues
// ascending in w, orthonormal eigenvectors in the columns of a (overwritten):
// dsyev_vu with the workspace the reference's factorizations pass.
void sym_eig(int n, double* a, double* w);
// Singular value decomposition a = U diag(s) V^T (square n x n, column-major):
// s descending; u = U, vt = V^T (both column-major): dgesvd_aa on a copy of a.
void svd_square(int n, const double* a, double* s, double* u, double* vt);
// Complex a = U diag(s) V^H (square n x n, column-major): s descending, u = U,
// vt = V^H: zgesvd_aa on a copy of a.
void svd_square_complex(int n, const Zc* a, double* s, Zc* u, Zc* vt);

// ---- factorizations (ic2.cpp, ilu2.cpp) ----

struct FactorStats {
  int64_t nzu = 0;      // stored first-order block count
  int64_t nzr = 0;      // second-order block count (discarded at the end)
  double densu = 0.0;   // stored values / nza(A)
  double densr = 0.0;   // second-order values / nza(A)
  double seconds = 0.0;
};

// Second-order incomplete Cholesky of an upper-stored SPD block matrix
// (diagonal block first in each row). Returns U by rows with inverted
// diagonal blocks, A ~ U^T U.
BlockCsr ich2(const BlockCsr& upper, const SlvParam& p, FactorStats* st = nullptr);
// Second-order incomplete LU of L-by-columns / U-by-rows parts with identical
// patterns (combine_lu). Returns L by columns and U by rows with inverted
// diagonal blocks, A ~ L U.
void ilu2(const BlockCsr& al, const BlockCsr& au, const SlvParam& p, BlockCsr& l, BlockCsr& u,
          FactorStats* st = nullptr);

// Complex block matrices on the host: the pattern of a BlockCsr whose `a`
// holds interleaved (re, im) doubles, two per entry, with `bsa` counting
// complex entries (the value of entry k sits at a[2k], a[2k + 1]).
// Real and imaginary parts as two real matrices of one pattern, and back.
void split_complex(const BlockCsr& c, BlockCsr& re, BlockCsr& im);
BlockCsr merge_complex(const BlockCsr& re, const BlockCsr& im);
// CSMatrixCS::CombineLU (smatrixcs.cpp:1185-1345): the L part transposed
// without conjugation (TranspMtrNoConj) and both parts padded to the union
// pattern; every step moves values only, so it is combine_lu on the real and
// imaginary parts.
void combine_lu_complex(const BlockCsr& a, BlockCsr& l, BlockCsr& u);
// Second-order incomplete LU of a complex matrix (CSMatrixCS::Ilu2,
// ilu2cs.cpp:1213-1567) from complex L-by-columns / U-by-rows parts with one
// pattern (combine_lu_complex). Returns complex L by columns and U by rows
// with inverted diagonal blocks, A ~ L U.
void ilu2_complex(const BlockCsr& al, const BlockCsr& au, const SlvParam& p, BlockCsr& l, BlockCsr& u,
                  FactorStats* st = nullptr);

// ---- synthetic systems (generate.cpp) ----

struct CellParams {
  double radius = 0.3;     // inclusion radius (cell edge = 1)
  double contrast = 10.0;  // inclusion/matrix stiffness ratio
  double poisson = 0.3;    // Poisson ratio
  double shift = 0.0;      // inclusion centre offset along x (cell units)
};

// Config-0 family: 2D heat, nx*ny cells, block size bla, 5-point stencil,
// symmetric; returns the upper-stored matrix and nrhs right-hand sides.
void gen_heat2d(int nx, int ny, int bla, uint64_t seed, BlockCsr& upper, std::vector<double>& rhs);
// Config-1 family: 3D nonsymmetric (skew advection), n^3 cells, 7-point, full storage.
void gen_seepage3d(int nx, int ny, int nz, int bla, double skew, uint64_t seed, BlockCsr& a,
                   std::vector<double>& rhs);
// Homogenization-cell family: periodic elasticity cell with a spherical
// inclusion, block size 24 (3 components x 8 terms), 7-point, block-SPD
// (a sum of positive semidefinite face elements plus a small multiple of the
// local stiffness on the diagonal), upper storage, 6 macroscopic-strain
// right-hand sides. The random template depends on (m, template_seed) only;
// the physics parameters enter smoothly.
void gen_elastic_cell(int m, const CellParams& cp, uint64_t template_seed, BlockCsr& upper,
                      std::vector<double>& rhs6);

}  // namespace bcm
