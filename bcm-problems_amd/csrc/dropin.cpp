// Drop-in entry points: the reference's AbstractSymmSolver / AbstractSolver
// (solvers.cpp:5383-5586, :5591-5794), the complex AbstractSolver
// (:7103-7398) and the dispatchers they call (CSVector::SolveSymm
// solvers.cpp:4452-4610, CSVector::Solver :1568-1860, CSVectorC::Solver
// :2588-2830), with the Krylov phase on the GPU.
//
// Steps kept from the reference: callback assembly (rows sorted by column,
// blocks copied in sorted order), RCM profile ordering (ordtype > 0; the
// parallel solver's nested dissection, PartBinaryTreeND),
// symmetric / general reordering, IC2 (symmetric) or CombineLU + ILU2
// (general), ordering of the right-hand side and initial guess, the Krylov
// solve, inverse ordering and CSTORESOL per block.
#include "../../include/bcmhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>

#include "device/solver.hpp"
#include "host/bcm_host.hpp"
#include "internal.hpp"

namespace {

struct Callbacks {
  void* obj;
  bcm_cb_nsup fnsup;
  bcm_cb_sprnds fsprnds;
  bcm_cb_iaja fiaja;
  bcm_cb_blockrow fblockrow;
  bcm_cb_rhs frhs;
  bcm_cb_iniguess finiguess;
  bcm_cb_storesol fstoresol;
};

struct CCallbacks {
  void* obj;
  bcm_cb_nsup fnsup;
  bcm_cb_sprnds fsprnds;
  bcm_cb_iaja fiaja;
  bcm_cb_blockrow_c fblockrow;
  bcm_cb_rhs_c frhs;
  bcm_cb_iniguess_c finiguess;
  bcm_cb_storesol_c fstoresol;
};

// Assembles the block matrix exactly as AbstractSolver does: the callback
// returns each row's blocks in its own ja order; rows are re-sorted by column
// and blocks copied in the sorted order (solvers.cpp:5683-5736; the complex
// solver the same way, :7196-7268). `width` doubles per entry: 1 real, 2
// complex (interleaved re/im; bsa counts entries).
template <class RowFn>
bcm::BlockCsr assemble_rows(void* obj, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds, bcm_cb_iaja fiaja, RowFn&& row,
                            int width) {
  int nsup = 0;
  fnsup(obj, &nsup);
  if (nsup <= 0) throw bcm::Error(bcm::kErrArg, "callback reported no block rows");
  std::vector<int> sprnd(nsup + 1);
  fsprnds(obj, sprnd.data());
  for (int i = 0; i < nsup; ++i)
    if (sprnd[i + 1] <= sprnd[i]) throw bcm::Error(bcm::kErrArg, "callback block sizes must be positive");
  int* ia = nullptr;
  int* ja = nullptr;
  fiaja(obj, &ia, &ja);
  if (!ia || !ja) throw bcm::Error(bcm::kErrArg, "callback returned no sparsity");
  bcm::BlockCsr m;
  try {
    m.nsup = nsup;
    m.sprnd = sprnd;
    m.ia.assign(ia, ia + nsup + 1);
    m.ja.resize(m.ia[nsup]);
    m.bsa.resize(m.ia[nsup]);
    size_t rowmax = 0, total = 0;
    for (int i = 0; i < nsup; ++i) {
      size_t row = 0;
      for (int j = ia[i]; j < ia[i + 1]; ++j) {
        if (ja[j] < 0 || ja[j] >= nsup) throw bcm::Error(bcm::kErrArg, "callback column index out of range");
        row += (size_t)m.bla(i) * m.bla(ja[j]);
      }
      rowmax = std::max(rowmax, row);
      total += row;
    }
    m.a.resize(total * width);
    std::vector<double> arow(std::max<size_t>(rowmax * width, 1));
    std::vector<size_t> off;
    std::vector<int> idx;
    size_t nz = 0;
    for (int i = 0; i < nsup; ++i) {
      const int nzj = ia[i + 1] - ia[i];
      const int bi = m.bla(i);
      off.assign(nzj, 0);
      size_t pos = 0;
      for (int k = 0; k < nzj; ++k) {
        off[k] = pos;
        pos += (size_t)bi * m.bla(ja[ia[i] + k]);
      }
      row(i, arow.data());
      idx.resize(nzj);
      std::iota(idx.begin(), idx.end(), 0);
      std::sort(idx.begin(), idx.end(), [&](int x, int y) { return ja[ia[i] + x] < ja[ia[i] + y]; });
      for (int k = 0; k < nzj; ++k) {
        const int src = idx[k];
        const int col = ja[ia[i] + src];
        const size_t sz = (size_t)bi * m.bla(col);
        m.ja[ia[i] + k] = col;
        m.bsa[ia[i] + k] = (int64_t)nz;
        std::memcpy(&m.a[nz * width], &arow[off[src] * width], sz * width * sizeof(double));
        nz += sz;
      }
    }
  } catch (...) {
    delete[] ia;
    delete[] ja;
    throw;
  }
  delete[] ia;
  delete[] ja;
  m.set_blamx();
  return m;
}

bcm::BlockCsr assemble(const Callbacks& cb) {
  return assemble_rows(cb.obj, cb.fnsup, cb.fsprnds, cb.fiaja,
                       [&](int i, double* r) { cb.fblockrow(cb.obj, i, r); }, 1);
}

void gather_blocks(const Callbacks& cb, const bcm::BlockCsr& m, bcm_cb_rhs f, std::vector<double>& v) {
  v.assign(m.n(), 0.0);
  for (int i = 0; i < m.nsup; ++i) f(cb.obj, i, &v[m.sprnd[i]]);
}

// Device vector of the drop-in; its allocation and free wait for any stream
// capture in progress on another thread (the library's set-up lock).
struct DevVec {
  double* p = nullptr;
  explicit DevVec(size_t n) {
    bcm::SetupGuard g;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess)
      throw bcm::Error(bcm::kErrHip, "hipMalloc failed");
  }
  ~DevVec() {
    bcm::SetupGuard g;
    if (p) (void)hipFree(p);
  }
};

// Solver kind and stopping rule of the serial dispatchers (SolveSymm / Solver(RS),
// solvers.cpp:4452-4610, 1568-1860) or, with par, of the parallel one (Solver(tree),
// solvers.cpp:1865-2041: ittype 1 parallel Cg, 2 Lanczos, else BlockGMRES).
void pick_solver(bool symmetric, bool par, const bcm_params& prm, bcm::SolveOptions& opt) {
  if (par) {
    if (prm.ittype == 1) {
      opt.kind = bcm::kCgPar;  // cg.cpp:607-614: sttype 0 absolute, else relative to ||r0||
      opt.ittype = prm.sttype == 0 ? 1 : 0;
    } else if (prm.ittype == 2) {
      opt.kind = bcm::kLanczos;  // lanczos.cpp:797: the parallel form's relative test is against ||b||
      opt.ittype = prm.sttype == 0 ? 1 : prm.sttype == 2 ? 2 : 3;
    } else {
      opt.kind = bcm::kBlockGmres;  // gmres.cpp:742-750: sttype 0 absolute, else relative to ||r0||
      opt.ittype = prm.sttype == 0 ? 1 : 0;
      opt.check_every = 1;  // a block GMRES step costs far more than a host poll of the done flag
    }
    return;
  }
  if (symmetric) {
    // Cg's stopping rules (cg.cpp:238-241): 0 relative to ||r0||, 1 absolute,
    // 2 max_i |r_i| <= |eps| (the unit norm vector of solvers.cpp:5554), and 3
    // (no reference counterpart) relative to ||b||.
    if (prm.ittype < 0 || prm.ittype > 3) throw bcm::Error(bcm::kErrArg, "AbstractSymmSolver: ittype must be 0..3");
    opt.kind = bcm::kPcg;
    opt.ittype = prm.ittype;
    return;
  }
  // Solver(RS) (solvers.cpp:1710-1800): ittype 1 two-sided Lanczos per
  // right-hand side; 2, 3, 4 the block Lanczos variants Central, Right, Left
  // (relative residual test every ichk iterations, lanczos.cpp:1536-1585).
  if (prm.ittype >= 2 && prm.ittype <= 4) {
    opt.kind = prm.ittype == 2 ? bcm::kBlzCentral : prm.ittype == 3 ? bcm::kBlzRight : bcm::kBlzLeft;
    opt.ichk = std::max(1, prm.ichk);
    return;
  }
  if (prm.ittype != 1) throw bcm::Error(bcm::kErrArg, "AbstractSolver: ittype must be 1 (Lanczos) or 2-4 (block Lanczos)");
  opt.kind = bcm::kLanczos;
  // Lanczos sttype (lanczos.cpp:263): 0 absolute, 1 relative to ||r0||, 2 max-norm.
  opt.ittype = prm.sttype == 1 ? 0 : prm.sttype == 2 ? 2 : 1;
}

int run(bool symmetric, bool par, bcm_params* p, const Callbacks& cb) {
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
      throw bcm::Error(bcm::kErrNoDevice, "no HIP device: the solve phase runs on the GPU only");
    bcm_params defaults;
    bcm_params_default(&defaults);
    bcm_params& prm = p ? *p : defaults;
    bcm::SlvParam slv;
    slv.ordtype = prm.ordtype;
    slv.memory = prm.memory;
    slv.tau1 = prm.tau1;
    slv.tau2 = prm.tau2;
    slv.theta = prm.theta;
    slv.sclmin = prm.sclmin;
    slv.pivmin = prm.pivmin;

    bcm::BlockCsr a = assemble(cb);
    if (symmetric)
      for (int i = 0; i < a.nsup; ++i)
        for (int j = a.ia[i]; j < a.ia[i + 1]; ++j)
          if (a.ja[j] < i) throw bcm::Error(bcm::kErrArg, "symmetric solver expects upper-triangular sparsity");
    a.validate(symmetric);
    std::vector<double> b, x;
    gather_blocks(cb, a, cb.frhs, b);
    gather_blocks(cb, a, cb.finiguess, x);

    // ordering: the parallel solver always dissects (PartBinaryTreeND,
    // solvers.cpp:5890: METIS_NodeND's layout, here this library's own
    // dissection); otherwise ordtype > 0 -> RCM profile ordering, 0 -> identity
    std::vector<int> order(a.nsup);
    const bool reorder = par || prm.ordtype != 0;
    if (par) order = bcm::nd_order(a);
    else if (prm.ordtype != 0) order = bcm::rcm_order(a);
    else std::iota(order.begin(), order.end(), 0);
    bcm::Matrix mat;
    mat.symmetric = symmetric;
    mat.h = reorder ? (symmetric ? bcm::ord_mtr_symm(a, order) : bcm::ord_mtr(a, order)) : a;
    std::vector<int> sp_new = mat.h.sprnd;
    if (reorder) {
      bcm::ord_vect('D', a.sprnd, sp_new, order, b.data(), a.n(), 1);
      bcm::ord_vect('D', a.sprnd, sp_new, order, x.data(), a.n(), 1);
    }
    bcm::Factor f;
    if (symmetric) {
      f.kind = 0;
      f.u = bcm::ich2(mat.h, slv, &f.stats);
    } else {
      bcm::BlockCsr al, au;
      bcm::combine_lu(mat.h, al, au);
      f.kind = 1;
      bcm::ilu2(al, au, slv, f.l, f.u, &f.stats);
    }
    prm.timefct = f.stats.seconds;
    prm.density = f.stats.densu;
    prm.density2 = f.stats.densr;

    bcm::Context ctx;
    const size_t n = a.n();
    DevVec db(n), dx(n);
    BCM_HIP(hipMemcpy(db.p, b.data(), n * 8, hipMemcpyHostToDevice));
    BCM_HIP(hipMemcpy(dx.p, x.data(), n * 8, hipMemcpyHostToDevice));
    bcm::SolveOptions opt;
    opt.niter = prm.niter;
    opt.eps = prm.eps;
    pick_solver(symmetric, par, prm, opt);
    opt.max_restarts = 0;
    bcm::SolveStats st;
    bcm::solve(mat, f, ctx, opt, db.p, dx.p, st);
    BCM_HIP(hipMemcpy(x.data(), dx.p, n * 8, hipMemcpyDeviceToHost));
    // serial Cg reports its loop index at exit, one past the last solution update (cg.cpp:209)
    prm.nitperf = opt.kind == bcm::kPcg ? st.iterations + 1 : st.iterations;
    // Cg reports errmin (cg.cpp:210): max_i |r_i| under ittype 2; the other
    // rules report the FP64 true relative residual here (the reference leaves 0)
    prm.resfin = opt.kind == bcm::kPcg && opt.ittype == 2 ? st.errmin : st.relres_true;
    prm.timeitr = st.seconds;
    if (reorder) bcm::ord_vect('I', a.sprnd, sp_new, order, x.data(), a.n(), 1);
    for (int i = 0; i < a.nsup; ++i) cb.fstoresol(cb.obj, i, &x[a.sprnd[i]]);
    if (st.breakdown) throw bcm::Error(bcm::kErrBreakdown, "Krylov breakdown (best iterate stored)");
    return BCM_OK;
  } catch (const bcm::Error& e) {
    bcm::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bcm::set_last_error(e.what());
    return BCM_E_INTERNAL;
  }
}

// The complex AbstractSolver (solvers.cpp:7103-7398) and CSVectorC::Solver
// (solvers.cpp:2588-2830): the complex matrix, factors and vectors are held
// as their real equivalents (complex.cpp), and the complex block GMRES of one
// right-hand side b is the real block GMRES of the two columns [b, i b] from
// [x0, i x0]: the real span of the complex Krylov space of b is the block
// Krylov space of those columns, so both minimize over the same set and the
// first column's iterate is the complex one after every step.
int run_complex(bcm_params* p, const CCallbacks& cb) {
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
      throw bcm::Error(bcm::kErrNoDevice, "no HIP device: the solve phase runs on the GPU only");
    bcm_params defaults;
    bcm_params_default(&defaults);
    bcm_params& prm = p ? *p : defaults;
    // ordtype = nparts of PartOrdMtr (solvers.cpp:2633, 2651-2657): one part is the
    // RCM profile ordering; more parts (METIS) and ordtype < 0 (OrderBrdMtr) are not restated
    if (prm.ordtype < 0 || prm.ordtype > 1)
      throw bcm::Error(bcm::kErrArg, "complex AbstractSolver: ordtype must be 0 (none) or 1 (RCM profile)");
    bcm::SlvParam slv;
    slv.ordtype = prm.ordtype;
    slv.memory = prm.memory;
    slv.tau1 = prm.tau1;
    slv.tau2 = prm.tau2;
    slv.theta = prm.theta;
    slv.sclmin = prm.sclmin;
    slv.pivmin = prm.pivmin;
    slv.fcttyp = prm.fcttyp;

    bcm::BlockCsr ac = assemble_rows(cb.obj, cb.fnsup, cb.fsprnds, cb.fiaja,
                                     [&](int i, double* r) { cb.fblockrow(cb.obj, i, r); }, 2);
    const int nsup = ac.nsup;
    const int n = ac.n();
    std::vector<double> bz(2 * (size_t)n, 0.0), xz(2 * (size_t)n, 0.0);
    for (int i = 0; i < nsup; ++i) {
      cb.frhs(cb.obj, i, &bz[2 * (size_t)ac.sprnd[i]]);
      cb.finiguess(cb.obj, i, &xz[2 * (size_t)ac.sprnd[i]]);
    }
    auto store = [&] {
      for (int i = 0; i < nsup; ++i) cb.fstoresol(cb.obj, i, &xz[2 * (size_t)ac.sprnd[i]]);
    };
    double bnorm2 = 0.0;
    for (double v : bz) bnorm2 += v * v;
    // no nonzero right-hand side: the guess is stored as it came, the intent
    // of the reference's column list (solvers.cpp:2725-2790), whose solver
    // then ends the process in the closing ZGESVD of an empty system
    if (bnorm2 == 0.0) {
      prm.nitperf = 0;
      prm.resfin = 0.0;
      store();
      return BCM_OK;
    }

    bcm::BlockCsr re, im;
    bcm::split_complex(ac, re, im);
    re.validate(false);
    std::vector<int> order(nsup);
    const bool reorder = prm.ordtype > 0;
    if (reorder) order = bcm::rcm_order(re);  // PartOrdMtr with one part: OrderPrfMtr(1) (ordmatr.cpp:1009-1017)
    else std::iota(order.begin(), order.end(), 0);
    bcm::BlockCsr co = reorder ? bcm::merge_complex(bcm::ord_mtr(re, order), bcm::ord_mtr(im, order)) : ac;
    const std::vector<int>& sp_new = co.sprnd;
    std::vector<int> sp2_old(nsup + 1), sp2_new(nsup + 1);  // block offsets in doubles of a complex vector
    for (int i = 0; i <= nsup; ++i) {
      sp2_old[i] = 2 * ac.sprnd[i];
      sp2_new[i] = 2 * sp_new[i];
    }
    if (reorder) {
      bcm::ord_vect('D', sp2_old, sp2_new, order, bz.data(), 2 * n, 1);
      bcm::ord_vect('D', sp2_old, sp2_new, order, xz.data(), 2 * n, 1);
    }
    bcm::BlockCsr al, au, l, u;
    bcm::combine_lu_complex(co, al, au);
    bcm::Factor f;
    f.kind = 1;
    f.cplx = true;
    bcm::ilu2_complex(al, au, slv, l, u, &f.stats);
    f.l = bcm::real_equivalent(nsup, sp_new.data(), l.ia.data(), l.ja.data(), l.a.data(), true);
    f.u = bcm::real_equivalent(nsup, sp_new.data(), u.ia.data(), u.ja.data(), u.a.data(), false);
    bcm::Matrix mat;
    mat.symmetric = false;
    mat.cplx = true;
    mat.h = bcm::real_equivalent(nsup, sp_new.data(), co.ia.data(), co.ja.data(), co.a.data(), false);
    prm.timefct = f.stats.seconds;
    prm.density = f.stats.densu;
    prm.density2 = f.stats.densr;

    // columns [b, i b] and [x0, i x0] in the real-equivalent layout
    const size_t n2 = 2 * (size_t)n;
    std::vector<double> rot(n2), breal(2 * n2), xreal(2 * n2);
    auto times_i = [&](const std::vector<double>& z) {
      for (size_t k = 0; k < (size_t)n; ++k) {
        rot[2 * k] = -z[2 * k + 1];
        rot[2 * k + 1] = z[2 * k];
      }
      return rot.data();
    };
    bcm::complex_to_real(nsup, sp_new.data(), bz.data(), breal.data(), 1);
    bcm::complex_to_real(nsup, sp_new.data(), times_i(bz), breal.data() + n2, 1);
    bcm::complex_to_real(nsup, sp_new.data(), xz.data(), xreal.data(), 1);
    bcm::complex_to_real(nsup, sp_new.data(), times_i(xz), xreal.data() + n2, 1);

    bcm::Context ctx;
    DevVec db(2 * n2), dx(2 * n2);
    BCM_HIP(hipMemcpy(db.p, breal.data(), 2 * n2 * 8, hipMemcpyHostToDevice));
    BCM_HIP(hipMemcpy(dx.p, xreal.data(), 2 * n2 * 8, hipMemcpyHostToDevice));
    bcm::SolveOptions opt;
    opt.kind = bcm::kBlockGmres;
    opt.niter = prm.niter;
    opt.eps = prm.eps;
    opt.ittype = 0;  // gmres.cpp:1308-1313: every column's ||r|| < eps ||r0||
    opt.max_restarts = 0;
    opt.check_every = 1;  // a block GMRES step costs far more than a host poll of the done flag
    bcm::SolveStats st;
    double resi[2] = {0.0, 0.0};
    auto t0 = std::chrono::steady_clock::now();
    bcm::solve_block_gmres(mat, f, ctx, opt, 2, db.p, dx.p, st, resi, nullptr);
    BCM_HIP(hipStreamSynchronize(ctx.stream));
    ctx.check_sweep_error();
    prm.timeitr = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    BCM_HIP(hipMemcpy(xreal.data(), dx.p, n2 * 8, hipMemcpyDeviceToHost));
    bcm::real_to_complex(nsup, sp_new.data(), xreal.data(), xz.data(), 1);
    prm.nitperf = st.iterations;
    prm.resfin = resi[0] / std::sqrt(bnorm2);
    if (reorder) bcm::ord_vect('I', sp2_old, sp2_new, order, xz.data(), 2 * n, 1);
    store();
    if (st.breakdown) throw bcm::Error(bcm::kErrBreakdown, "Krylov breakdown (best iterate stored)");
    return BCM_OK;
  } catch (const bcm::Error& e) {
    bcm::set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    bcm::set_last_error(e.what());
    return BCM_E_INTERNAL;
  }
}

}  // namespace

extern "C" {

int bcm_abstract_symm_solver(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                             bcm_cb_iaja fiaja, bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs,
                             bcm_cb_iniguess finiguess, bcm_cb_storesol fstoresol) {
  if (!fnsup || !fsprnds || !fiaja || !fblockrow || !frhs || !finiguess || !fstoresol) return BCM_E_ARG;
  return run(true, false, p, Callbacks{pgener, fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol});
}

int bcm_abstract_solver(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds, bcm_cb_iaja fiaja,
                        bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs, bcm_cb_iniguess finiguess,
                        bcm_cb_storesol fstoresol) {
  if (!fnsup || !fsprnds || !fiaja || !fblockrow || !frhs || !finiguess || !fstoresol) return BCM_E_ARG;
  return run(false, false, p, Callbacks{pgener, fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol});
}

int bcm_abstract_par_solver(void* pgener, void* comm, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                            bcm_cb_iaja fiaja, bcm_cb_blockrow fblockrow, bcm_cb_rhs frhs,
                            bcm_cb_iniguess finiguess, bcm_cb_storesol fstoresol) {
  (void)comm;  // one rank: the whole system lives on this process's GPU
  if (!fnsup || !fsprnds || !fiaja || !fblockrow || !frhs || !finiguess || !fstoresol) return BCM_E_ARG;
  return run(false, true, p, Callbacks{pgener, fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol});
}

int bcm_abstract_solver_c(void* pgener, bcm_params* p, bcm_cb_nsup fnsup, bcm_cb_sprnds fsprnds,
                          bcm_cb_iaja fiaja, bcm_cb_blockrow_c fblockrow, bcm_cb_rhs_c frhs,
                          bcm_cb_iniguess_c finiguess, bcm_cb_storesol_c fstoresol) {
  if (!fnsup || !fsprnds || !fiaja || !fblockrow || !frhs || !finiguess || !fstoresol) return BCM_E_ARG;
  return run_complex(p, CCallbacks{pgener, fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol});
}

}  // extern "C"
