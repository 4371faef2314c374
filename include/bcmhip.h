/*
 * libbcmhip: MI355X (gfx950) solve phase of Bslib's block preconditioned
 * Krylov library. Plain C ABI: opaque handles, plain pointers and sizes, int
 * status codes (BCM_OK == 0). Device pointers are HIP device allocations
 * holding column-major n x nrhs fp64 vectors.
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the Bslib / Bcmlib sources).
 */
#ifndef BCMHIP_H
#define BCMHIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference throws C strings instead) ---- */
#define BCM_OK 0
#define BCM_E_ARG 1
#define BCM_E_NOT_SPD 2       /* ich2rs.cpp:118-123, :575 negative eigenvalue / pivot */
#define BCM_E_WORKSPACE 3     /* ich2rs.cpp:1246-1250, ilu2rs.cpp:1655-1663 */
#define BCM_E_BREAKDOWN 4     /* cg.cpp:270-278, lanczos.cpp:797-799 */
#define BCM_E_HIP 5
#define BCM_E_NOCONV 6
#define BCM_E_NODEVICE 7
#define BCM_E_INTERNAL 8

/* ---- solver kinds and precisions ---- */
#define BCM_PCG 0        /* CSVector::Cg serial form, cg.cpp:31-349 */
#define BCM_BICGSTAB 1   /* no reference counterpart (nonsymmetric drop-in) */
#define BCM_FGMRES 2     /* flexible GMRES(m); reference block GMRES is gmres.cpp:337 */
#define BCM_LANCZOS 3    /* CSVector::Lanczos, lanczos.cpp:34-371 */
#define BCM_CGPAR 4      /* CSVector::Cg parallel form with operator plug-ins, cg.cpp:355-771 */
#define BCM_BLOCKPCG 5   /* batched PCG over many cells and right-hand sides (bcm_batch_solve) */
#define BCM_BLOCKGMRES 6 /* CSVector::BlockGMRES, gmres.cpp:337-1038 (bcm_solve_block; bcm_solve: 1 column) */
/* Block Lanczos (block Golub-Kahan bidiagonalization with block Givens
 * rotations) of Solver(RS), solvers.cpp:1798-1800, by preconditioning side
 * (bcm_solve_block; bcm_solve: 1 column): */
#define BCM_BLZ_CENTRAL 7 /* CSVector::BlockLanczosCentral, lanczos.cpp:2467-3225 (ittype 2) */
#define BCM_BLZ_RIGHT 8   /* CSVector::BlockLanczosRight, lanczos.cpp:932-1718 (ittype 3) */
#define BCM_BLZ_LEFT 9    /* CSVector::BlockLanczosLeft, lanczos.cpp:1719-2466 (ittype 4) */

#define BCM_F64 0
#define BCM_F32 1
#define BCM_BF16 2

typedef struct bcm_matrix_s* bcm_matrix;   /* CSMatrixRS, smatrix.h:1040-1048 */
typedef struct bcm_factor_s* bcm_factor;   /* IC2 U / ILU2 L,U factors (CSMatrixRS) */
typedef struct bcm_context_s* bcm_context; /* one HIP stream + solver workspace */

/* Solver parameters: the solve-path subset of CSlvParam (slvparam.h:19-113,
 * defaults slvparam.h:121-214). Fields after `eps` are outputs. */
typedef struct bcm_params {
  int msglev;
  int ordtype;    /* 1: RCM profile ordering (OrderPrfMtr), 0: none */
  double memory;  /* factor workspace in units of nza(A); <= 0: unbounded */
  double tau1, tau2, theta, sclmin, pivmin;
  int ittype;     /* Cg stopping rule (cg.cpp:238-241): 0 ||r|| <= eps ||r0||, 1 ||r|| <= eps,
                     2 max_i |r_i| <= |eps| (MaxProdSVect with a unit norm vector), 3 ||r|| <= eps ||b|| */
  int sttype;
  int niter;
  int ichk;
  double eps;
  int nitperf;    /* out */
  double resfin;  /* out: final true relative residual */
  double timefct; /* out: host factorization seconds */
  double timeitr; /* out: solve seconds */
  double density, density2; /* out: factor fill (first / second order) per nza(A) */
  int fcttyp;     /* complex ILU2 only (ilu2cs.cpp:613-651, 1014-1049): 0 keeps exactly the blocks of
                     the pattern of A, otherwise tau1 / tau2 thresholds (slvparam.h:33; default 1) */
} bcm_params;

typedef struct bcm_solve_opts {
  int kind;          /* BCM_PCG, BCM_BICGSTAB, BCM_FGMRES, BCM_LANCZOS, BCM_CGPAR, BCM_BLOCKGMRES */
  int niter;
  double eps;
  int ittype;        /* 0 ||r|| <= eps ||r0||, 1 ||r|| <= eps, 2 (BCM_PCG only) max_i |r_i| <= |eps|,
                        3 ||b - A x|| <= eps ||b|| */
  int check_every;   /* device state polled every k iterations */
  int restart;       /* FGMRES cycle length */
  int precision;     /* preconditioner storage/application: BCM_F64, BCM_F32, BCM_BF16 */
  int max_restarts;  /* true-residual restarts */
  int ichk;          /* block Lanczos: iterations between residual tests (CSlvParam::ichk, slvparam.h:174) */
} bcm_solve_opts;

typedef struct bcm_stats {
  int iterations, restarts, breakdown, converged;
  double relres_true, relres_rec, seconds;
  double errmin;  /* BCM_PCG with ittype 2: max_i |r_i| at the last check (Cg's errmin, cg.cpp:194-210) */
} bcm_stats;

/* ---- library ---- */
const char* bcm_version(void);
const char* bcm_last_error(void);            /* message of the last failed call on this thread */
void bcm_params_default(bcm_params* p);      /* CSlvParam::CSlvParam, slvparam.h:121-214 */
void bcm_solve_opts_default(bcm_solve_opts* o);

/* ---- host block matrices (CSMatrixRS) ---- */
/* Copies nsup block rows: sprnd[nsup+1] block-size prefix sums, ia[nsup+1],
 * ja[nzja] (ascending per row, diagonal first when symmetric), a: blocks in ja
 * order, each column-major bla_i x bla_j (mvm.cpp:2560). symmetric != 0 means
 * upper-triangular storage as AbstractSymmSolver assembles it (solvers.cpp:5383). */
int bcm_matrix_create(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, int symmetric,
                      bcm_matrix* out);
int bcm_matrix_destroy(bcm_matrix m);
/* Complex block matrix (CSMatrixCS, smatrix.h:1499; its products MvmA,
 * mvmc.cpp:160) held as its real equivalent: each complex block C of
 * b_i x b_j becomes the real block [[Re C, -Im C], [Im C, Re C]] of
 * 2b_i x 2b_j, so every entry point taking a bcm_matrix (products, solvers)
 * applies the complex operator to vectors in the real-equivalent layout
 * (bcm_complex_to_real). a: nza complex values interleaved (re, im) -- the
 * reference's dcmplx array -- blocks column-major in ja order. Products agree
 * with the reference's complex arithmetic to rounding. */
int bcm_cmatrix_create(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, bcm_matrix* out);
/* Complex ILU2 factors as CSMatrixCS::Ilu2 returns them (ilu2cs.cpp; L by
 * columns with transposed blocks, U by rows, inverted diagonal blocks first in
 * every row; values interleaved re/im), held as their real equivalents for the
 * sweeps of SolveL / SolveU (mvmc.cpp:1723, 2506) and every preconditioned
 * solver; bcm_cfactor_ilu2 computes them. */
int bcm_cfactor_create(int nsup, const int* sprnd, const int* lia, const int* lja, const double* la,
                       const int* uia, const int* uja, const double* ua, bcm_factor* out);
/* CSMatrixCS::CombineLU (smatrixcs.cpp:1185) + CSMatrixCS::Ilu2
 * (ilu2cs.cpp:1213-1567) of a complex matrix in the given order (arrays as
 * bcm_cmatrix_create; the diagonal block first in every row), bit for bit
 * the reference's factors (its ZGESVD restated from LAPACK 3.0), held as
 * their real equivalents like bcm_cfactor_create's. Reads tau1, tau2, theta,
 * sclmin, pivmin, memory and fcttyp; writes timefct, density, density2. */
int bcm_cfactor_ilu2(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, bcm_params* p,
                     bcm_factor* out);
/* Complex vectors (interleaved re/im, n x nrhs, CSVectorC's layout) to and
 * from the real-equivalent layout (per block row [Re z; Im z], 2n x nrhs). */
int bcm_complex_to_real(int nsup, const int* sprnd, const double* z, double* x, int nrhs);
int bcm_real_to_complex(int nsup, const int* sprnd, const double* x, double* z, int nrhs);
int bcm_matrix_info(bcm_matrix m, int* nsup, int* n, int* nzja, long long* nza, int* symmetric);
int bcm_matrix_export(bcm_matrix m, int* sprnd, int* ia, int* ja, double* a);

/* ---- seeded synthetic systems (stand-ins for the Bcmlib producers) ---- */
int bcm_gen_heat2d(int nx, int ny, int bla, unsigned long long seed, bcm_matrix* upper, double* rhs);
int bcm_gen_seepage3d(int nx, int ny, int nz, int bla, double skew, unsigned long long seed, bcm_matrix* a,
                      double* rhs);
int bcm_gen_elastic_cell(int m, double radius, double contrast, double poisson, double shift,
                         unsigned long long seed, bcm_matrix* upper, double* rhs6);

/* ---- ordering (ordmatr.cpp:28-227 OrderPrfMtr(1) -> rcm.cpp genrcm) ---- */
int bcm_rcm_order(bcm_matrix m, int* order);                       /* order[old] = new */
/* CSMatrix::PartBinaryTreeND (ordmatr.cpp:1698-1805) in one process: the
 * nested-dissection ordering of the block pattern (this library's own
 * dissection in place of METIS_NodeND, with its layout: parts, then their
 * separator) and the blocks of the binary dissection tree over nproc virtual
 * ranks (a power of two), as FindSeparator and CTree::PartitionTree derive
 * them (ordmatr.cpp:1204-1396, tree.cpp:2414). order[old] = new (nsup
 * entries); blks gets nblks + 1 entries (at most 2 nproc), blk2cpu nblks. */
int bcm_part_binary_tree_nd(bcm_matrix m, int nproc, int* order, int* nblks, int* blks, int* blk2cpu);
/* CSMatrix::FindSeparator (smatrix.cpp:4355-4488) on a symmetric pattern of
 * n rows: ind1 = first row of the second part, ind2 = first separator row
 * (-1, -1 when the pattern has no such split). */
int bcm_find_separator(int n, const int* ia, const int* ja, int* ind1, int* ind2);
/* OrdMtrSymm (smatrixrs.cpp:456) for symmetric, OrdMtr (smatrixrs.cpp:345) otherwise. */
int bcm_matrix_reorder(bcm_matrix m, const int* order, bcm_matrix* out);
/* CSVector::OrdVect (svector.cpp:418): type 'D' applies order, 'I' undoes it. */
int bcm_vector_reorder(bcm_matrix original, const int* order, char type, double* v, int nrhs);

/* ---- binary files (the reference's WriteMtrBin / WriteVectBin layout) ---- */
/* CSMatrixRS::WriteMtrBin (smatrixrs.cpp:3210-3227): raw ints m, n, nsupr,
 * nsupc, nlist, nzja, nza, blamx, sprndr, sprndc, list, ia, ja, bsa, then the
 * doubles a. */
int bcm_matrix_write_bin(bcm_matrix m, const char* path);
/* Reads that layout (square, every block row listed, blocks in ja order);
 * symmetric as in bcm_matrix_create. The reference writes this layout but has
 * no block-matrix reader; CSMatrixR::ReadMtrBin (smatrixr.cpp:3852) is its
 * point-matrix counterpart. */
int bcm_matrix_read_bin(const char* path, int symmetric, bcm_matrix* out);
/* CSVector::WriteVectBin (svector.cpp:537-546) of an n x nrhs host vector. */
int bcm_vector_write_bin(const char* path, int n, int nrhs, const double* v);
/* CSVector::ReadVectBin (svector.cpp:488-507): *n and *nrhs always, the values
 * into v (n * nrhs doubles) when v is not NULL. */
int bcm_vector_read_bin(const char* path, int* n, int* nrhs, double* v);

/* ---- dense kernels of the factorizations ---- */
/* dsyev_("V", "U") as the IC2 calls it (lapackd.cpp:865-878 -> LAPACK 3.0
 * DSYEV, ich2rs.cpp:110): a (n x n column-major, upper triangle read) is
 * overwritten by the eigenvectors, w receives the eigenvalues ascending. */
int bcm_dsyev(int n, double* a, double* w);
/* dgesvd_("A", "A") on a square matrix as the ILU2 calls it (lapackd.cpp:883-905
 * -> LAPACK 3.0 DGESVD, ilu2rs.cpp:126): a = U diag(s) VT, s descending; a is
 * overwritten. */
int bcm_dgesvd(int n, double* a, double* s, double* u, double* vt);
/* zgesvd_("A", "A") of a square complex matrix as the complex ILU2 calls it
 * (ilu2cs.cpp:132, 725; LAPACK 3.0 ZGESVD): a = U diag(s) VT with VT = V^H,
 * complex arrays interleaved re/im, column-major; a is overwritten; n >= 2
 * (the factorization treats 1 x 1 blocks itself). */
int bcm_zgesvd(int n, double* a, double* s, double* u, double* vt);

/* CSMatrixRS::CombineLU (smatrixrs.cpp:1645-1803): splits a full-storage
 * matrix into L stored by columns (as the transpose of its strict lower part,
 * diagonal blocks transposed) and U by rows, both padded with zero blocks to
 * the union of the two patterns: the input pair of bcm_factor_ilu2. */
int bcm_combine_lu(bcm_matrix a, bcm_matrix* lower, bcm_matrix* upper);

/* ---- host factorizations ---- */
/* CSMatrixRS::Ich2 (ich2rs.cpp:1075) of an upper-stored SPD matrix. */
int bcm_factor_ic2(bcm_matrix upper, bcm_params* p, bcm_factor* out);
/* CombineLU (smatrixrs.cpp:1645) + CSMatrixRS::Ilu2 (ilu2rs.cpp:1458). */
int bcm_factor_ilu2(bcm_matrix a, bcm_params* p, bcm_factor* out);
int bcm_factor_destroy(bcm_factor f);
/* which: 0 = L (by columns; for IC2 the U factor), 1 = U (by rows). */
int bcm_factor_info(bcm_factor f, int which, int* nzja, long long* nza, int* levels_fwd, int* levels_bwd);
int bcm_factor_export(bcm_factor f, int which, int* ia, int* ja, double* a);

/* ---- device memory (HIP allocations owned by the caller) ---- */
int bcm_device_count(int* count);
int bcm_dev_alloc(unsigned long long bytes, void** out);
int bcm_dev_free(void* p);
int bcm_copy_to_device(void* dst, const void* src, unsigned long long bytes);
int bcm_copy_to_host(void* dst, const void* src, unsigned long long bytes);
int bcm_dev_zero(void* p, unsigned long long bytes);
/* Waits for the device, then frees what the library keeps between calls for
 * the life of the process (the pinned host staging buffers of matrix and
 * factor uploads). Call it after destroying every handle and before the
 * process exits, so that no HIP resource of the library is left to the
 * runtime's own teardown; later calls allocate again on demand. Not while a
 * build runs on another thread. No counterpart in the reference. */
int bcm_shutdown(void);

/* ---- device context ---- */
int bcm_context_create(bcm_context* out);
int bcm_context_destroy(bcm_context c);
int bcm_context_sync(bcm_context c);
/* Triangular sweep schedule: 0 = one launch per dependency level;
 * 1 = one sync-free cooperative launch per sweep (default): resident waves
 * own positions of the rows sorted by dependency level round-robin and wait
 * (bounded) on the values of the rows they depend on; 2 = as 1 over the
 * natural sweep order. The schedules agree to rounding: per row they add
 * the same products, grouped differently across lanes for some block sizes. */
int bcm_context_set_sweep_mode(bcm_context c, int mode);
/* Resident waves per CU of the sync-free sweeps (1..64, default 8; fewer
 * waves than block rows make each wave own several rows in turn). */
int bcm_context_set_sweep_waves(bcm_context c, int waves_per_cu);
/* Test hook: marks the sync-free sweeps' dependency wait as having reached its
 * bound, so that every wait of the next solve exits at once and the solve
 * takes its recovery path (repeated from x0 with cooperative launches). No
 * counterpart in the reference. */
int bcm_context_raise_wait_flag(bcm_context c);

/* Per-kernel-class timing with HIP events on the context's stream (the
 * reference's clock() statistics per operator, cg.cpp:743-764). Times are
 * the sum over launches; bytes are the algorithmic bytes of those launches. */
typedef struct bcm_profile {
  double spmv_ms, sweep_ms;
  long long spmv_launches, sweep_launches;
  double spmv_bytes, sweep_bytes;
} bcm_profile;
int bcm_context_profile(bcm_context c, int enable);
int bcm_context_profile_read(bcm_context c, bcm_profile* out, int reset);

/* ---- device operators: the CMVM / CSOLVEL / CSOLVEU / CSOLVELU plug-ins (svector.h:34-41) ---- */
/* y = A x: CSMatrixRS::MvmA (mvm.cpp:2530); symmetric matrices: MvmASymm (mvm.cpp:3008). */
int bcm_mvm(bcm_context c, bcm_matrix a, const double* x, double* y, int nrhs);
/* y = A^T x: CSMatrixRS::MvmAt (mvm.cpp:2593). */
int bcm_mvm_t(bcm_context c, bcm_matrix a, const double* x, double* y, int nrhs);
/* r = b - A x (operator- after MvmA, svector.cpp:285). */
int bcm_residual(bcm_context c, bcm_matrix a, const double* b, const double* x, double* r, int nrhs);
/* Forward sweep of CSMatrixRS::SolveL (mvm.cpp:3140) on factor part `which`. */
int bcm_solve_l(bcm_context c, bcm_factor f, int which, const double* x, double* y, int nrhs);
/* Backward sweep of CSMatrixRS::SolveU (mvm.cpp:3410) on factor part `which`. */
int bcm_solve_u(bcm_context c, bcm_factor f, int which, const double* x, double* y, int nrhs);
/* z = M^-1 r (CParSolveLU::SolveLU, gsmatrix.h:1008-1071) in the given precision; transpose: M^-T. */
int bcm_precond(bcm_context c, bcm_factor f, const double* r, double* z, int nrhs, int precision, int transpose);
/* No counterpart in the reference: bcm_precond for callers that cannot
 * allocate device memory (code under HIP stream capture, serving loops that
 * forbid hipMalloc). Same result as bcm_precond, bit for bit; allocates
 * nothing: the vector between the two sweeps lives in a fixed device array
 * of 32 MiB (plus 256 bytes of control words) per GPU, kept until the
 * process ends, so n * nrhs must not
 * exceed 4,194,304 (BCM_E_ARG otherwise). The factor's device image for this
 * precision and direction must exist already (one bcm_precond call with the
 * same precision and transpose flag builds it; BCM_E_ARG otherwise). The
 * sweeps follow the context's sweep mode; the sync-free modes keep their
 * control words in the same array, cleared by a kernel on every call. Every
 * call on one GPU shares the array: calls that overlap on one GPU, from any
 * stream or context, race. r and z must not alias. */
int bcm_precond_noalloc(bcm_context c, bcm_factor f, const double* r, double* z, int nrhs, int precision,
                        int transpose);
/* ScProdSVect (svector.cpp:364): result written to host memory. */
int bcm_dot(bcm_context c, int n, const double* x, const double* y, double* result);

/* ---- GPU solve: x (device) holds the initial guess on entry ---- */
int bcm_solve(bcm_context c, bcm_matrix a, bcm_factor f, const bcm_solve_opts* o, const double* b, double* x,
              bcm_stats* st);
/* Block GMRES (o->kind BCM_BLOCKGMRES) or block Lanczos (BCM_BLZ_CENTRAL /
 * RIGHT / LEFT) over nrhs (1..8) right-hand sides. Block GMRES, replacing
 * CSVector::BlockGMRES (gmres.cpp:337-1038) with the operator pair of
 * CParMvmA::MvmA / CParSolveLU::SolveLU (gsmatrix.h:758-817, 1008-1071):
 * b and x (device) are n x nrhs column-major like CSVector's vect, x holds
 * x0; o->niter is the Krylov block dimension (no restart); o->ittype 0 tests
 * ||r_c|| < eps ||r0_c|| (sttype != 0), 1 tests ||r_c|| < eps (sttype 0),
 * 3 tests ||r_c|| < eps ||b_c||. Columns with r0 = 0 are left as they are.
 * resi (host, nrhs, may be NULL) receives ||b_c - A x_c||; hist (host,
 * niter x nrhs, may be NULL) the recursion's residual norms after each step.
 * Block Lanczos (lanczos.cpp:932-3225): o->niter caps the iterations, every
 * o->ichk iterations each column's true residual is tested against eps
 * times its initial norm (Right: also ||Ahat^T r||), the solve ends when all
 * pass; hist is not written. Columns with r0 = 0 are left as they are. */
int bcm_solve_block(bcm_context c, bcm_matrix a, bcm_factor f, const bcm_solve_opts* o, int nrhs, const double* b,
                    double* x, bcm_stats* st, double* resi, double* hist);

/* ---- batched solves: many independent cells (block-diagonal union on the
 *      device), nrhs right-hand sides each; one launch sequence for every
 *      (cell, right-hand side) pair. Block counterpart of the reference's
 *      per-system Cg (cg.cpp:31-349) as the homogenization drivers call it
 *      once per cell and load case. ---- */
typedef struct bcm_batch_s* bcm_batch;
typedef struct bcm_batch_stats {
  int pairs, converged, breakdowns, iters_max, restarts;
  long long iters_sum;
  double relres_max, seconds;
  int recovered; /* 1: a sync-free dependency wait reached its bound; the solve was repeated from x0 with
                    level-scheduled sweeps (which never wait) and these are the repeat's results */
} bcm_batch_stats;
typedef struct bcm_batch_tune {
  int sweep_mode;    /* 0 level-scheduled sweeps, 1 sync-free sweeps */
  int waves_per_cu;  /* sync-free sweeps: resident waves per CU */
  int level_order;   /* sync-free sweeps: 1 rows in dependency-level order, 0 natural order */
  int graph;         /* 1: replay check_every iterations as one captured hipGraph */
  int check_every;   /* iterations between host polls */
  double ms_per_iter;
} bcm_batch_tune;
/* Builds the union of ncell matrices and their IC2 factors (one block size).
 * prec_mask: bit p keeps the factor sweeps in precision p (bit 0 F64, 1 F32, 2 BF16).
 * The per-cell host work runs on up to 16 host threads of its own; the copies
 * go through pinned staging buffers (32 MB each, at most one per thread) that
 * the library keeps for reuse until the process ends. */
int bcm_batch_create(int ncell, const bcm_matrix* mats, const bcm_factor* facts, int prec_mask, bcm_batch* out);
int bcm_batch_destroy(bcm_batch b);
/* Union length n and the element offset of every cell (elem0: ncell + 1 entries, may be NULL). */
int bcm_batch_layout(bcm_batch b, int* ncell, long long* n, long long* elem0);
/* Block PCG for all pairs; rhs and x are device arrays of n x nrhs (column q
 * holds every cell's q-th right-hand side at the cell's offsets); x holds x0.
 * o->precision selects the factor sweeps (F32/BF16: flexible CG with an FP64
 * outer recurrence). fallback, bit 0 (1): a pair with ||b - A x0|| >= ||b||
 * restarts from x0 = 0; bit 1 (2, no reference counterpart): first the nrhs
 * guesses X of every cell are replaced by their A-optimal combination X C,
 * (X^T A X) C = X^T B -- each column no farther from its solution in the
 * A-norm than its own guess or than zero -- at the cost of three vector
 * passes (a cell whose guesses are dependent to rounding keeps them); other
 * values BCM_E_ARG. iters/relres: host arrays of ncell * nrhs entries
 * (cell-major) or NULL. nrhs is 1, 3 or 6. */
int bcm_batch_solve(bcm_context c, bcm_batch b, const bcm_solve_opts* o, int nrhs, const double* rhs, double* x,
                    int fallback, bcm_batch_stats* st, int* iters, double* relres);
/* Times the candidate schedules on this batch and keeps the fastest;
 * allow_graph = 0 leaves hipGraph replay out of the candidates. */
int bcm_batch_autotune(bcm_context c, bcm_batch b, const bcm_solve_opts* o, int nrhs, const double* rhs,
                       int allow_graph, bcm_batch_tune* out);
int bcm_batch_get_tune(bcm_batch b, bcm_batch_tune* t);
int bcm_batch_set_tune(bcm_batch b, const bcm_batch_tune* t);

/* ---- learned initial guess (no counterpart in the reference) ----
 * MLP nin -> nh (tanh) -> nh (tanh) -> nout with BF16 weights and FP32
 * accumulation on the matrix cores. Parameters are one FP32 array
 * [W1 b1 W2 b2 W3 b3], W_l row-major fan_in x fan_out. Activations are
 * row-major batch x features, device memory. */
typedef struct bcm_mlp_s* bcm_mlp;
int bcm_mlp_create(int nin, int nhidden, int nout, unsigned long long seed, bcm_mlp* out);
int bcm_mlp_destroy(bcm_mlp m);
int bcm_mlp_nparam(bcm_mlp m, long long* nparam);
int bcm_mlp_set_params(bcm_mlp m, const float* host);  /* also resets the Adam state */
int bcm_mlp_get_params(bcm_mlp m, float* host);
int bcm_mlp_get_grads(bcm_mlp m, float* host);
int bcm_mlp_forward(bcm_context c, bcm_mlp m, int batch, const float* x, float* y);
/* gradients of mean((MLP(x) - t)^2) (kept in the model); *loss (host) may be NULL */
int bcm_mlp_backward(bcm_context c, bcm_mlp m, int batch, const float* x, const float* t, double* loss);
int bcm_mlp_adam_step(bcm_context c, bcm_mlp m, float lr);

/* Initial-guess model for a family of cells with n unknowns and nrhs
 * right-hand sides: solved cells are stored as snapshots, fit() builds
 * nmodes POD modes per right-hand side (randomized range finder on the
 * device) and trains the MLP (nparam -> nhidden -> nhidden -> nrhs*nmodes)
 * on the snapshots' coefficients; predict() writes x0 for new cells. Cell
 * vectors use the batch layout: column q holds the cells' q-th vectors one
 * after another (n each). */
typedef struct bcm_guess_s* bcm_guess;
typedef struct bcm_guess_report {
  int snapshots, modes, steps;
  double loss_first, loss_final, pod_energy, seconds_pod, seconds_train;
} bcm_guess_report;
int bcm_guess_create(long long n, int nrhs, int nparam, int nmodes, int nhidden, int capacity,
                     unsigned long long seed, bcm_guess* out);
int bcm_guess_destroy(bcm_guess g);
int bcm_guess_add(bcm_context c, bcm_guess g, int count, const double* params, const double* x);
int bcm_guess_fit(bcm_context c, bcm_guess g, int steps, double lr, bcm_guess_report* rep);
int bcm_guess_predict(bcm_context c, bcm_guess g, int count, const double* params, double* x0);

/* ---- drop-in entry points (callback ABI of smatrix.h:15-21; C++ references
 *      are passed as pointers, which is the same machine-level convention) ---- */
typedef void (*bcm_cb_nsup)(void* obj, int* nsup);
typedef void (*bcm_cb_sprnds)(void* obj, int* sprnds);
typedef void (*bcm_cb_iaja)(void* obj, int** ia, int** ja);  /* callee allocates with new[], library frees with delete[] */
typedef void (*bcm_cb_blockrow)(void* obj, int isup, double* arow);
typedef void (*bcm_cb_rhs)(void* obj, int isup, double* rhs);
typedef void (*bcm_cb_iniguess)(void* obj, int isup, double* xini);
typedef void (*bcm_cb_storesol)(void* obj, int isup, double* sol);
/* CBLOCKROWC / CRHSC / CINIGUESSC / CSTORESOLC (smatrix.h:21-24): complex
 * data, interleaved (re, im) doubles, behind a void pointer. */
typedef void (*bcm_cb_blockrow_c)(void* obj, int isup, void* arow);
typedef void (*bcm_cb_rhs_c)(void* obj, int isup, void* rhs);
typedef void (*bcm_cb_iniguess_c)(void* obj, int isup, void* xini);
typedef void (*bcm_cb_storesol_c)(void* obj, int isup, void* sol);

/* AbstractSymmSolver (solvers.cpp:5383): assemble upper storage, RCM, IC2,
 * GPU PCG, return the solution through storesol. */
int bcm_abstract_symm_solver(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                             bcm_cb_iaja fiaja, bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs,
                             bcm_cb_iniguess finiguess, bcm_cb_storesol fstoresol);
/* AbstractSolver (solvers.cpp:5591): assemble, RCM, ILU2, GPU Krylov as
 * Solver(RS) dispatches (solvers.cpp:1710-1800): ittype 1 two-sided Lanczos,
 * 2 / 3 / 4 block Lanczos Central / Right / Left; other values BCM_E_ARG. */
int bcm_abstract_solver(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds, bcm_cb_iaja fiaja,
                        bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs, bcm_cb_iniguess finiguess,
                        bcm_cb_storesol fstoresol);

/* AbstractParSolver (solvers.cpp:5799-6224) on one rank: assemble, nested
 * dissection (PartBinaryTreeND, solvers.cpp:5890, with this library's own
 * dissection in place of METIS_NodeND; on one rank the tree is a single
 * node), ILU2 of the dissected matrix (what Ilu2Schur reduces to on one rank
 * is not restated: the serial Ilu2 is used), then the kind of
 * Solver(tree) (solvers.cpp:1865-2041): ittype 1 parallel-form Cg, 2
 * two-sided Lanczos (relative test against ||b||, lanczos.cpp:797), otherwise
 * block GMRES. comm (the reference's CMPIComm*) is not used: the system lives
 * on this process's GPU. */
int bcm_abstract_par_solver(void* pgener, void* comm, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                            bcm_cb_iaja fiaja, bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs,
                            bcm_cb_iniguess finiguess, bcm_cb_storesol fstoresol);

/* The complex AbstractSolver (solvers.cpp:7103-7398) and the dispatcher it
 * calls, CSVectorC::Solver (solvers.cpp:2588-2830): assemble the complex
 * block rows, RCM profile ordering when ordtype = 1 (PartOrdMtr with one
 * part is OrderPrfMtr(1)), none when 0 (ordtype > 1, METIS parts, and
 * ordtype < 0, OrderBrdMtr: BCM_E_ARG),
 * CombineLU + complex ILU2 (bcm_cfactor_ilu2), then block GMRES
 * (CSVectorC::BlockGMRES, gmres.cpp:1044-1466: one cycle of niter steps,
 * stopped when ||r|| < eps ||r0||) on the GPU, and the solution through
 * fstoresol. A zero right-hand side stores the initial guess unchanged and
 * returns BCM_OK: the reference's solver skips zero columns
 * (solvers.cpp:2725-2790), but its closing diagnostic ZGESVD of the empty
 * system (gmres.cpp:1405) then fails its LWORK check and LAPACK's XERBLA
 * stops the process. The
 * complex Krylov space over the real equivalent: the real block GMRES of
 * the columns [b, i b] from [x0, i x0] spans it step for step, so the
 * iterate after k steps is the complex block GMRES iterate. Writes nitperf,
 * resfin (FP64 true relative residual), timefct, timeitr, density, density2. */
int bcm_abstract_solver_c(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                          bcm_cb_iaja fiaja, bcm_cb_blockrow_c fblockrow, bcm_cb_rhs_c frhs,
                          bcm_cb_iniguess_c finiguess, bcm_cb_storesol_c fstoresol);

/* Callback memory for bcm_cb_iaja (operator new[] of nint ints, as new_struct in Bcmlib/utils.h:29-35). */
int* bcm_alloc_ints(int nint);

#ifdef __cplusplus
}
#endif

#endif /* BCMHIP_H */
