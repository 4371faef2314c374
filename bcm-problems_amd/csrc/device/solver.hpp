// Handles behind the C ABI: matrices, factors and the per-stream solver
// context, plus the GPU-resident Krylov drivers.
#pragma once

#include <memory>

#include "device.hpp"

namespace bcm {

// A block matrix as the caller sees it (host copy, reference layout) plus
// its device images. A symmetric matrix is kept upper-stored on the host
// (as the reference's AbstractSymmSolver assembles it) and expanded to full
// storage on the device.
struct Matrix {
  BlockCsr h;
  bool symmetric = false;
  bool cplx = false;  // h is the real equivalent of a complex matrix: products read its complex blocks
  DevBsr d;       // full matrix on device
  DevBsr dt;      // transpose (two-sided Lanczos only)
  bool has_d = false, has_dt = false;
  void to_device(bool with_transpose);
};

// IC2 / ILU2 factor: host storage in the reference layout and the device
// sweeps of M^-1 (fwd then bwd) and, on request, of M^-T (fwdt then bwdt).
struct Factor {
  int kind = 0;            // 0 IC2: A ~ U^T U (U only); 1 ILU2: A ~ L U
  bool cplx = false;       // l, u are real equivalents of complex factors: FP64 sweeps read their complex blocks
  BlockCsr l, u;           // L by columns, U by rows, inverted diagonal blocks
  FactorStats stats;
  std::vector<int> order;  // ordering the factor was computed in (order[old] = new), may be empty
  // One device image per storage precision (fp64, fp32, bf16), built on first use.
  struct Image {
    DevTri fwd, bwd, fwdt, bwdt;
    bool has = false, has_t = false;
  };
  Image img[3];
  const Image& to_device(int prec, bool with_transpose);
  const BlockCsr& lower() const { return kind == 0 ? u : l; }
};

// Scalars and flags of one running solver, in device memory.
struct KState {
  double sc[32];
  int st[8];  // [0] done, [1] iteration, [2] breakdown, [3] converged-on-s (BiCGStab)
};

// Per-kernel-class timing with HIP events recorded on the context's stream.
enum ProfKind : int { kProfSpmv = 0, kProfSweep = 1, kProfKinds = 2 };
struct ProfAcc {
  double ms = 0.0;
  long long launches = 0;
  double bytes = 0.0;
};

struct Context {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  RedWork rw;
  int sweep_mode = 1;   // 0: one launch per dependency level, 1: one sync-free launch per sweep
  SweepSync sync;
  SweepSync* sweep_sync() { return sweep_mode == 1 ? &sync : nullptr; }
  void check_sweep_error();
  DevBuf kstate;        // KState
  KState* host_state = nullptr;  // pinned mirror
  std::vector<DevBuf> vecs;      // solver work vectors
  Context();
  ~Context();
  KState* dstate() const { return kstate.as<KState>(); }
  double* vec(int i, size_t nbytes);  // returns work vector i of at least nbytes
  void pull_state();                  // copy state to the pinned mirror and wait
  // profiling
  bool prof = false;
  ProfAcc prof_acc[kProfKinds];
  std::vector<hipEvent_t> ev_pool;
  struct Pending { int kind; hipEvent_t e0, e1; double bytes; };
  std::vector<Pending> pending;
  hipEvent_t prof_begin();
  void prof_end(int kind, hipEvent_t e0, double bytes);
  void prof_flush();
  // Bytes of launches whose work is only known on the device (row-gated
  // products and sweeps): accumulated by device kernels per kind, added to
  // prof_acc when read.
  DevBuf prof_dev;
  double* prof_dev_bytes();  // kProfKinds device doubles (zeroed on first use)
  void prof_collect();       // adds and clears the device-side bytes
};

// Times the launches issued during its lifetime as one kernel of `kind`.
struct ProfScope {
  Context& c;
  int kind;
  double bytes;
  hipEvent_t e0 = nullptr;
  // row-gated launch: bytes = sum over parts of active[c] * part_bytes[c] (device)
  const int* gate_active = nullptr;
  const double* part_bytes = nullptr;
  int nparts = 0;
  ProfScope(Context& ctx, int k, double b) : c(ctx), kind(k), bytes(b) {
    if (c.prof) e0 = c.prof_begin();
  }
  ProfScope(Context& ctx, int k, const int* active, const double* pbytes, int np)
      : c(ctx), kind(k), bytes(0.0), gate_active(active), part_bytes(pbytes), nparts(np) {
    if (c.prof) e0 = c.prof_begin();
  }
  ~ProfScope();
};

inline double spmv_bytes(const DevBsr& A, int nrhs) {
  return (double)A.value_bytes() + (double)A.index_bytes() + 16.0 * A.n * nrhs;
}
inline double sweep_bytes(const DevTri& T, int nrhs) {
  return (double)T.value_bytes() + (double)T.index_bytes() + 2.0 * prec_bytes(T.prec == kF64 ? kF64 : kF32) * T.n * nrhs;
}

enum SolverKind : int {
  kPcg = 0,          // reference serial Cg (cg.cpp:31-349)
  kBicgstab = 1,     // right-preconditioned BiCGStab (no reference counterpart)
  kFgmres = 2,       // flexible GMRES(m), mixed-precision preconditioner
  kLanczos = 3,      // reference two-sided Lanczos (lanczos.cpp:34-371)
  kCgPar = 4,        // reference parallel-form Cg (cg.cpp:355-771)
  kBlockPcg = 5,     // block PCG over several right-hand sides
  kBlockGmres = 6,   // reference block GMRES (gmres.cpp:337-1038)
  kBlzCentral = 7,   // reference block Lanczos, central preconditioning (lanczos.cpp:2467-3225)
  kBlzRight = 8,     // right preconditioning (lanczos.cpp:932-1718)
  kBlzLeft = 9,      // left preconditioning (lanczos.cpp:1719-2466)
};

// Block Lanczos variant of a solver kind: 0 Right, 1 Left, 2 Central, -1 none.
inline int blz_variant(int kind) {
  return kind == kBlzRight ? 0 : kind == kBlzLeft ? 1 : kind == kBlzCentral ? 2 : -1;
}

struct SolveOptions {
  int kind = kPcg;
  int niter = 1000;
  double eps = 1e-10;
  int ittype = 3;        // 0: ||r|| <= eps ||r0||, 1: ||r|| <= eps, 2 (PCG): max|r_i| <= |eps|, 3: ||r|| <= eps ||b||
  int check_every = 8;   // host polls the device state every k iterations
  int restart = 30;      // FGMRES cycle length
  int prec = kF64;       // precision of the preconditioner application
  int max_restarts = 3;  // true-residual restarts when the recursion drifted
  int ichk = 5;          // block Lanczos: iterations between residual tests (slvparam.h:174)
};

struct SolveStats {
  int iterations = 0;
  int restarts = 0;
  int breakdown = 0;
  int converged = 0;
  double relres_true = 0.0;   // ||b - A x|| / ||b|| recomputed in fp64 at the end
  double relres_rec = 0.0;    // from the recursion
  double seconds = 0.0;
  int launches = 0;
  double errmin = 0.0;        // PCG with ittype 2: max|r_i| at the last check (Cg's errmin, cg.cpp:210)
};

// Solves A x = b for one right-hand side (b, x device pointers, x holds x0).
void solve(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
           SolveStats& st);

// Applies M^-1 (or M^-T) in the given precision: z = M^-1 r.
void precond_apply(Factor& F, Context& ctx, const double* r, double* z, int prec, bool transpose, int nrhs,
                   const int* skip = nullptr);

// precond_apply without any allocation (noalloc.hip): the vector between the
// sweeps lives in a fixed per-device array of precond_noalloc_max_elems()
// doubles; the factor image must already be on the device.
void precond_noalloc(Factor& F, Context& ctx, const double* r, double* z, int prec, bool transpose, int nrhs);
size_t precond_noalloc_max_elems();

// z = M^-1 r through the sweep pair (t1 then t2) in the given precision
// (fp64 vectors, or fp32 vectors converted on the way in and out).
void precond_sweeps(const DevTri& t1, const DevTri& t2, Context& ctx, const double* r, double* z, int prec, int nrhs,
                    const int* skip = nullptr, RowGate gate = RowGate());

// FGMRES(m) driver (fgmres.hip).
void solve_fgmres(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
                  SolveStats& st);

// ---------------------------------------------------------------------------
// Batched solves: many independent cells held as one block-diagonal union on
// the device, nrhs right-hand sides per cell, and one launch sequence for all
// (cell, right-hand side) pairs, each pair with its own recurrence scalars.

// Schedule of a batched solve, picked by batch_tune.
struct BatchTune {
  int sweep_mode = 1;     // 0: level-scheduled sweeps, 1: sync-free sweeps
  int waves_per_cu = 8;   // sync-free sweeps: resident waves per CU
  int level_order = 1;    // sync-free sweeps: rows in dependency-level order (0: natural order)
  int graph = 0;          // 1: replay check_every iterations as one captured hipGraph
  int check_every = 8;    // iterations between host polls of the done flag
  double ms_per_iter = 0.0;
};

struct BatchGraph;  // captured iteration block (bpcg.hip)

struct Batch {
  int ncell = 0, b = 0, kind = 0;
  int64_t n = 0;                      // union length (elements per right-hand side)
  std::vector<int64_t> elem0;         // ncell + 1 element offsets of the cells
  DevBsr A;                           // union of the cell matrices (full storage, fp64)
  Factor::Image img[3];               // union sweeps per factor precision
  int nchunks = 0;                    // reduction chunks (each inside one cell)
  std::vector<int> cell_chunk0;       // ncell + 1
  DevBuf chunk_cell, chunk_lo, chunk_hi, dcell_chunk0;
  DevBuf row_cell;                    // cell of every block row of the union
  // per cell: block rows, blocks of A, off-diagonal blocks of the forward
  // and backward sweeps (the algorithmic bytes of row-gated launches)
  std::vector<int64_t> cell_rows, cell_ablk, cell_fdep, cell_bdep;
  BatchTune tune;
  bool tuned = false;
  std::unique_ptr<BatchGraph> graph;
  Batch();
  ~Batch();
  void drop_graph();  // after a schedule change
};

struct BatchStats {
  int pairs = 0, converged = 0, breakdowns = 0, iters_max = 0, restarts = 0;
  long long iters_sum = 0;
  double relres_max = 0.0, seconds = 0.0;
  int recovered = 0;  // the solve was repeated with level-scheduled sweeps after a wait reached its bound
};

// Builds the device union of the cells' matrices and factors (all IC2 or all
// ILU2, one block size). prec_mask: bit p uploads the factor sweeps in
// precision p (fp64 = bit 0, fp32 = bit 1, bf16 = bit 2).
void batch_build(Batch& bt, const std::vector<const Matrix*>& mats, const std::vector<const Factor*>& facts,
                 int prec_mask);
// Block PCG over all pairs. b, x: device, union layout, nrhs columns of
// length bt.n; x holds x0. With opt_fallback, a pair whose ||b - A x0|| >=
// ||b|| restarts from x0 = 0. iters/relres (host, ncell * nrhs, cell-major)
// may be null.
// fallback: pairs whose initial residual is not below ||b|| restart from 0;
// galerkin: first replace each cell's nrhs guesses by their A-optimal
// combination (X C with X^T A X C = X^T B).
void batch_solve(Batch& bt, Context& ctx, const SolveOptions& opt, int nrhs, const double* b, double* x,
                 bool fallback, BatchStats& st, int* iters, double* relres, bool galerkin = false);
// Times the candidate schedules on this batch (a few iterations each, on
// scratch copies of b) and keeps the fastest in bt.tune.
BatchTune batch_tune(Batch& bt, Context& ctx, const SolveOptions& opt, int nrhs, const double* b,
                     bool allow_graph = true);

// Two-sided Lanczos driver (lanczos.hip).
void solve_lanczos(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, const double* b, double* x,
                   SolveStats& st);

// Block Lanczos driver (blanczos.hip), variant 0 Right, 1 Left, 2 Central,
// over nrhs (1..8) right-hand sides (n x nrhs column-major, x holds x0);
// resi (host, nrhs, optional) receives ||b_c - A x_c||.
void solve_block_lanczos(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int variant, int nrhs,
                         const double* b, double* x, SolveStats& st, double* resi = nullptr);

// Block GMRES driver (bgmres.hip) over nrhs (1..8) right-hand sides: b and x
// are n x nrhs column-major, x holds x0; opt.niter is the number of block
// steps (no restart). resi (host, nrhs) receives ||b_c - A x_c||; hist (host,
// niter x nrhs, optional) the recursion's residual norms after each step.
void solve_block_gmres(Matrix& A, Factor& F, Context& ctx, const SolveOptions& opt, int nrhs, const double* b,
                       double* x, SolveStats& st, double* resi = nullptr, double* hist = nullptr);

}  // namespace bcm
