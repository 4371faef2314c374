// C ABI of libbcmhip (include/bcmhip.h). Every entry point converts
// exceptions into status codes and keeps the message for bcm_last_error().
#include "../../include/bcmhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "device/learn.hpp"
#include "device/solver.hpp"
#include "host/bcm_host.hpp"
#include "internal.hpp"

struct bcm_matrix_s {
  bcm::Matrix m;
};
struct bcm_factor_s {
  bcm::Factor f;
};
struct bcm_context_s {
  bcm::Context c;
};
struct bcm_batch_s {
  bcm::Batch b;
};
struct bcm_mlp_s {
  std::unique_ptr<bcm::Mlp> m;
};
struct bcm_guess_s {
  std::unique_ptr<bcm::Guess> g;
};

namespace {

thread_local std::string g_err;

template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return BCM_OK;
  } catch (const bcm::Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "out of host memory";
    return BCM_E_WORKSPACE;
  } catch (const std::exception& e) {
    g_err = e.what();
    return BCM_E_INTERNAL;
  } catch (const char* s) {
    g_err = s;
    return BCM_E_INTERNAL;
  }
}

void need(bool ok, const char* what) {
  if (!ok) throw bcm::Error(bcm::kErrArg, what);
}

void need_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    throw bcm::Error(bcm::kErrNoDevice, "no HIP device: the solve phase runs on the GPU only");
}

bcm::SlvParam to_slv(const bcm_params* p) {
  bcm::SlvParam s;
  if (!p) return s;
  s.msglev = p->msglev;
  s.ordtype = p->ordtype;
  s.memory = p->memory;
  s.tau1 = p->tau1;
  s.tau2 = p->tau2;
  s.theta = p->theta;
  s.sclmin = p->sclmin;
  s.pivmin = p->pivmin;
  s.ittype = p->ittype;
  s.sttype = p->sttype;
  s.niter = p->niter;
  s.ichk = p->ichk;
  s.eps = p->eps;
  s.fcttyp = p->fcttyp;
  return s;
}

bcm::SolveOptions to_opts(const bcm_solve_opts* o) {
  bcm::SolveOptions s;
  if (!o) return s;
  s.kind = o->kind;
  s.niter = o->niter;
  s.eps = o->eps;
  s.ittype = o->ittype;
  s.check_every = std::max(1, o->check_every);
  s.restart = std::max(1, o->restart);
  s.prec = o->precision;
  s.max_restarts = std::max(0, o->max_restarts);
  s.ichk = std::max(1, o->ichk);
  return s;
}

void fill_stats(const bcm::SolveStats& s, bcm_stats* st) {
  if (!st) return;
  st->iterations = s.iterations;
  st->restarts = s.restarts;
  st->breakdown = s.breakdown;
  st->converged = s.converged;
  st->relres_true = s.relres_true;
  st->relres_rec = s.relres_rec;
  st->seconds = s.seconds;
  st->errmin = s.errmin;
}

bcm::BlockCsr make_csr(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a) {
  need(nsup > 0 && sprnd && ia && ja && a, "matrix: null argument or empty matrix");
  bcm::BlockCsr m;
  m.nsup = nsup;
  m.sprnd.assign(sprnd, sprnd + nsup + 1);
  m.ia.assign(ia, ia + nsup + 1);
  need(m.ia[0] == 0, "matrix: ia[0] must be 0");
  for (int i = 0; i < nsup; ++i) need(m.sprnd[i + 1] > m.sprnd[i], "matrix: empty block");
  const int nzja = m.ia[nsup];
  need(nzja >= 0, "matrix: bad ia");
  m.ja.assign(ja, ja + nzja);
  m.bsa.resize(nzja);
  int64_t off = 0;
  for (int i = 0; i < nsup; ++i) {
    need(m.ia[i + 1] >= m.ia[i], "matrix: ia not monotone");
    for (int j = m.ia[i]; j < m.ia[i + 1]; ++j) {
      need(m.ja[j] >= 0 && m.ja[j] < nsup, "matrix: column index out of range");
      m.bsa[j] = off;
      off += (int64_t)m.bla(i) * m.bla(m.ja[j]);
    }
  }
  m.a.assign(a, a + off);
  m.set_blamx();
  return m;
}

void build_factor_ic2(const bcm::BlockCsr& upper, const bcm::SlvParam& p, bcm::Factor& f) {
  f.kind = 0;
  f.u = bcm::ich2(upper, p, &f.stats);
}

void build_factor_ilu2(const bcm::BlockCsr& a, const bcm::SlvParam& p, bcm::Factor& f) {
  bcm::BlockCsr al, au;
  bcm::combine_lu(a, al, au);
  f.kind = 1;
  bcm::ilu2(al, au, p, f.l, f.u, &f.stats);
}

const bcm::BlockCsr& factor_part(const bcm::Factor& f, int which) { return which == 0 ? f.lower() : f.u; }

}  // namespace

void bcm::set_last_error(const std::string& msg) { g_err = msg; }

extern "C" {

const char* bcm_version(void) { return "libbcmhip 0.1 (gfx950)"; }

const char* bcm_last_error(void) { return g_err.c_str(); }

void bcm_params_default(bcm_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  bcm::SlvParam s;
  p->msglev = s.msglev;
  p->ordtype = s.ordtype;
  p->memory = s.memory;
  p->tau1 = s.tau1;
  p->tau2 = s.tau2;
  p->theta = s.theta;
  p->sclmin = s.sclmin;
  p->pivmin = s.pivmin;
  p->ittype = s.ittype;
  p->sttype = s.sttype;
  p->niter = s.niter;
  p->ichk = s.ichk;
  p->eps = s.eps;
  p->fcttyp = s.fcttyp;
}

void bcm_solve_opts_default(bcm_solve_opts* o) {
  if (!o) return;
  bcm::SolveOptions s;
  o->kind = s.kind;
  o->niter = s.niter;
  o->eps = s.eps;
  o->ittype = s.ittype;
  o->check_every = s.check_every;
  o->restart = s.restart;
  o->precision = s.prec;
  o->max_restarts = s.max_restarts;
  o->ichk = s.ichk;
}

int bcm_matrix_create(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, int symmetric,
                      bcm_matrix* out) {
  return guarded([&] {
    need(out != nullptr, "matrix_create: out is null");
    auto* h = new bcm_matrix_s;
    try {
      h->m.h = make_csr(nsup, sprnd, ia, ja, a);
      h->m.symmetric = symmetric != 0;
      h->m.h.validate(symmetric != 0);
      if (symmetric)
        for (int i = 0; i < nsup; ++i)
          for (int j = ia[i]; j < ia[i + 1]; ++j) need(ja[j] >= i, "symmetric matrix must be upper-stored");
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int bcm_cmatrix_create(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a,
                       bcm_matrix* out) {
  return guarded([&] {
    need(out && sprnd && ia && ja && a && nsup >= 0, "cmatrix_create: bad argument");
    auto* h = new bcm_matrix_s;
    try {
      h->m.h = bcm::real_equivalent(nsup, sprnd, ia, ja, a, false);
      h->m.symmetric = false;
      h->m.cplx = true;
      h->m.h.validate(false);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int bcm_cfactor_create(int nsup, const int* sprnd, const int* lia, const int* lja, const double* la,
                       const int* uia, const int* uja, const double* ua, bcm_factor* out) {
  return guarded([&] {
    need(out && sprnd && lia && lja && la && uia && uja && ua && nsup >= 0, "cfactor_create: bad argument");
    auto* h = new bcm_factor_s;
    try {
      h->f.kind = 1;
      h->f.cplx = true;
      h->f.l = bcm::real_equivalent(nsup, sprnd, lia, lja, la, true);
      h->f.u = bcm::real_equivalent(nsup, sprnd, uia, uja, ua, false);
      for (int i = 0; i < nsup; ++i) {
        need(h->f.l.ia[i] < h->f.l.ia[i + 1] && h->f.l.ja[h->f.l.ia[i]] == i, "cfactor_create: L row without its diagonal block first");
        need(h->f.u.ia[i] < h->f.u.ia[i + 1] && h->f.u.ja[h->f.u.ia[i]] == i, "cfactor_create: U row without its diagonal block first");
      }
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int bcm_cfactor_ilu2(int nsup, const int* sprnd, const int* ia, const int* ja, const double* a, bcm_params* p,
                     bcm_factor* out) {
  return guarded([&] {
    need(out && sprnd && ia && ja && a && nsup > 0, "cfactor_ilu2: bad argument");
    bcm::BlockCsr c;
    c.nsup = nsup;
    c.sprnd.assign(sprnd, sprnd + nsup + 1);
    c.ia.assign(ia, ia + nsup + 1);
    need(c.ia[0] == 0, "cfactor_ilu2: ia[0] must be 0");
    c.ja.assign(ja, ja + c.ia[nsup]);
    c.bsa.resize(c.ia[nsup]);
    int64_t nz = 0;
    for (int i = 0; i < nsup; ++i) {
      need(c.bla(i) > 0, "cfactor_ilu2: empty block row");
      for (int j = c.ia[i]; j < c.ia[i + 1]; ++j) {
        need(c.ja[j] >= 0 && c.ja[j] < nsup, "cfactor_ilu2: block column out of range");
        need(j == c.ia[i] || c.ja[j] > c.ja[j - 1], "cfactor_ilu2: block columns must ascend within a row");
        c.bsa[j] = nz;
        nz += (int64_t)c.bla(i) * c.bla(c.ja[j]);
      }
    }
    c.a.assign(a, a + 2 * nz);
    c.set_blamx();
    bcm::BlockCsr al, au, l, u;
    bcm::combine_lu_complex(c, al, au);
    bcm::FactorStats stats;
    bcm::ilu2_complex(al, au, to_slv(p), l, u, &stats);
    auto* h = new bcm_factor_s;
    try {
      h->f.kind = 1;
      h->f.cplx = true;
      h->f.l = bcm::real_equivalent(nsup, sprnd, l.ia.data(), l.ja.data(), l.a.data(), true);
      h->f.u = bcm::real_equivalent(nsup, sprnd, u.ia.data(), u.ja.data(), u.a.data(), false);
      h->f.stats = stats;
    } catch (...) {
      delete h;
      throw;
    }
    if (p) {
      p->timefct = stats.seconds;
      p->density = stats.densu;
      p->density2 = stats.densr;
    }
    *out = h;
  });
}

int bcm_complex_to_real(int nsup, const int* sprnd, const double* z, double* x, int nrhs) {
  return guarded([&] {
    need(sprnd && z && x && nsup >= 0 && nrhs >= 1, "complex_to_real: bad argument");
    bcm::complex_to_real(nsup, sprnd, z, x, nrhs);
  });
}

int bcm_real_to_complex(int nsup, const int* sprnd, const double* x, double* z, int nrhs) {
  return guarded([&] {
    need(sprnd && z && x && nsup >= 0 && nrhs >= 1, "real_to_complex: bad argument");
    bcm::real_to_complex(nsup, sprnd, x, z, nrhs);
  });
}

int bcm_matrix_destroy(bcm_matrix m) {
  return guarded([&] { delete m; });
}

int bcm_matrix_info(bcm_matrix m, int* nsup, int* n, int* nzja, long long* nza, int* symmetric) {
  return guarded([&] {
    need(m != nullptr, "matrix_info: null handle");
    if (nsup) *nsup = m->m.h.nsup;
    if (n) *n = m->m.h.n();
    if (nzja) *nzja = m->m.h.nzja();
    if (nza) *nza = (long long)m->m.h.nza();
    if (symmetric) *symmetric = m->m.symmetric ? 1 : 0;
  });
}

int bcm_matrix_export(bcm_matrix m, int* sprnd, int* ia, int* ja, double* a) {
  return guarded([&] {
    need(m != nullptr, "matrix_export: null handle");
    const auto& h = m->m.h;
    if (sprnd) std::copy(h.sprnd.begin(), h.sprnd.end(), sprnd);
    if (ia) std::copy(h.ia.begin(), h.ia.end(), ia);
    if (ja) std::copy(h.ja.begin(), h.ja.end(), ja);
    if (a) std::copy(h.a.begin(), h.a.end(), a);
  });
}

int bcm_gen_heat2d(int nx, int ny, int bla, unsigned long long seed, bcm_matrix* upper, double* rhs) {
  return guarded([&] {
    need(upper != nullptr, "gen_heat2d: out is null");
    auto* h = new bcm_matrix_s;
    std::vector<double> b;
    try {
      bcm::gen_heat2d(nx, ny, bla, seed, h->m.h, b);
    } catch (...) {
      delete h;
      throw;
    }
    h->m.symmetric = true;
    if (rhs) std::copy(b.begin(), b.end(), rhs);
    *upper = h;
  });
}

int bcm_gen_seepage3d(int nx, int ny, int nz, int bla, double skew, unsigned long long seed, bcm_matrix* a,
                      double* rhs) {
  return guarded([&] {
    need(a != nullptr, "gen_seepage3d: out is null");
    auto* h = new bcm_matrix_s;
    std::vector<double> b;
    try {
      bcm::gen_seepage3d(nx, ny, nz, bla, skew, seed, h->m.h, b);
    } catch (...) {
      delete h;
      throw;
    }
    h->m.symmetric = false;
    if (rhs) std::copy(b.begin(), b.end(), rhs);
    *a = h;
  });
}

int bcm_gen_elastic_cell(int m, double radius, double contrast, double poisson, double shift,
                         unsigned long long seed, bcm_matrix* upper, double* rhs6) {
  return guarded([&] {
    need(upper != nullptr, "gen_elastic_cell: out is null");
    auto* h = new bcm_matrix_s;
    std::vector<double> b;
    bcm::CellParams cp;
    cp.radius = radius;
    cp.contrast = contrast;
    cp.poisson = poisson;
    cp.shift = shift;
    try {
      bcm::gen_elastic_cell(m, cp, seed, h->m.h, b);
    } catch (...) {
      delete h;
      throw;
    }
    h->m.symmetric = true;
    if (rhs6) std::copy(b.begin(), b.end(), rhs6);
    *upper = h;
  });
}

int bcm_rcm_order(bcm_matrix m, int* order) {
  return guarded([&] {
    need(m && order, "rcm_order: null argument");
    auto o = bcm::rcm_order(m->m.h);
    std::copy(o.begin(), o.end(), order);
  });
}

int bcm_matrix_reorder(bcm_matrix m, const int* order, bcm_matrix* out) {
  return guarded([&] {
    need(m && order && out, "matrix_reorder: null argument");
    std::vector<int> o(order, order + m->m.h.nsup);
    std::vector<char> seen(o.size(), 0);
    for (int v : o) {
      need(v >= 0 && v < (int)o.size() && !seen[v], "matrix_reorder: order is not a permutation");
      seen[v] = 1;
    }
    auto* h = new bcm_matrix_s;
    try {
      h->m.symmetric = m->m.symmetric;
      h->m.h = m->m.symmetric ? bcm::ord_mtr_symm(m->m.h, o) : bcm::ord_mtr(m->m.h, o);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int bcm_part_binary_tree_nd(bcm_matrix m, int nproc, int* order, int* nblks, int* blks, int* blk2cpu) {
  return guarded([&] {
    need(m && order && nblks && blks && blk2cpu, "part_binary_tree_nd: null argument");
    need(nproc >= 1 && (nproc & (nproc - 1)) == 0, "part_binary_tree_nd: nproc must be a power of two");
    const std::vector<int> o = bcm::nd_order(m->m.h);
    const bcm::BlockCsr ordered = m->m.symmetric ? bcm::ord_mtr_symm(m->m.h, o) : bcm::ord_mtr(m->m.h, o);
    const bcm::NdTree t = bcm::part_binary_tree_nd(ordered, nproc);
    std::copy(o.begin(), o.end(), order);
    *nblks = (int)t.blk2cpu.size();
    std::copy(t.blks.begin(), t.blks.end(), blks);
    std::copy(t.blk2cpu.begin(), t.blk2cpu.end(), blk2cpu);
  });
}

int bcm_find_separator(int n, const int* ia, const int* ja, int* ind1, int* ind2) {
  return guarded([&] {
    need(n >= 0 && ia && ja && ind1 && ind2, "find_separator: bad argument");
    bcm::find_separator(n, ia, ja, *ind1, *ind2);
  });
}

int bcm_matrix_write_bin(bcm_matrix m, const char* path) {
  return guarded([&] {
    need(m && path, "matrix_write_bin: bad argument");
    bcm::write_mtr_bin(m->m.h, path);
  });
}

int bcm_matrix_read_bin(const char* path, int symmetric, bcm_matrix* out) {
  return guarded([&] {
    need(path && out, "matrix_read_bin: bad argument");
    auto* h = new bcm_matrix_s;
    try {
      h->m.h = bcm::read_mtr_bin(path);
      h->m.symmetric = symmetric != 0;
      h->m.h.validate(symmetric != 0);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int bcm_vector_write_bin(const char* path, int n, int nrhs, const double* v) {
  return guarded([&] {
    need(path && n >= 0 && nrhs >= 1 && (v || n == 0), "vector_write_bin: bad argument");
    bcm::write_vect_bin(path, n, nrhs, v);
  });
}

int bcm_vector_read_bin(const char* path, int* n, int* nrhs, double* v) {
  return guarded([&] {
    need(path && n && nrhs, "vector_read_bin: bad argument");
    std::vector<double> buf;
    bcm::read_vect_bin(path, n, nrhs, v ? &buf : nullptr);
    if (v && !buf.empty()) std::memcpy(v, buf.data(), buf.size() * sizeof(double));
  });
}

int bcm_dsyev(int n, double* a, double* w) {
  return guarded([&] {
    need(n >= 0 && (n == 0 || (a && w)), "dsyev: bad argument");
    bcm::sym_eig(n, a, w);
  });
}

int bcm_dgesvd(int n, double* a, double* s, double* u, double* vt) {
  return guarded([&] {
    need(n >= 0 && (n == 0 || (a && s && u && vt)), "dgesvd: bad argument");
    std::vector<double> work((size_t)10 * n + 1);
    const int info = bcm::lapack30::dgesvd_aa(n, a, n, s, u, n, vt, n, work.data(), 10 * n);
    if (info != 0) throw bcm::Error(bcm::kErrInternal, "Error in the Lapack routine DGESVD");
  });
}

int bcm_zgesvd(int n, double* a, double* s, double* u, double* vt) {
  return guarded([&] {
    need(n >= 2 && a && s && u && vt, "zgesvd: bad argument (n >= 2)");
    std::vector<bcm::Zc> work((size_t)10 * n + 1);
    std::vector<double> rwork((size_t)10 * n + 1);
    const int info = bcm::lapack30::zgesvd_aa(n, reinterpret_cast<bcm::Zc*>(a), n, s, reinterpret_cast<bcm::Zc*>(u),
                                              n, reinterpret_cast<bcm::Zc*>(vt), n, work.data(), 10 * n,
                                              rwork.data());
    if (info != 0) throw bcm::Error(bcm::kErrInternal, "Error in the Lapack routine ZGESVD");
  });
}

int bcm_combine_lu(bcm_matrix a, bcm_matrix* lower, bcm_matrix* upper) {
  return guarded([&] {
    need(a && lower && upper, "combine_lu: null argument");
    need(!a->m.symmetric, "combine_lu: expects a matrix in full storage");
    auto* l = new bcm_matrix_s;
    auto* u = new bcm_matrix_s;
    try {
      bcm::combine_lu(a->m.h, l->m.h, u->m.h);
    } catch (...) {
      delete l;
      delete u;
      throw;
    }
    *lower = l;
    *upper = u;
  });
}

int bcm_vector_reorder(bcm_matrix original, const int* order, char type, double* v, int nrhs) {
  return guarded([&] {
    need(original && order && v && nrhs > 0, "vector_reorder: null argument");
    need(type == 'D' || type == 'I', "vector_reorder: type must be 'D' or 'I'");
    const auto& h = original->m.h;
    std::vector<int> o(order, order + h.nsup);
    std::vector<int> spn(h.nsup + 1, 0);
    for (int i = 0; i < h.nsup; ++i) spn[o[i] + 1] = h.bla(i);
    for (int i = 0; i < h.nsup; ++i) spn[i + 1] += spn[i];
    bcm::ord_vect(type, h.sprnd, spn, o, v, h.n(), nrhs);
  });
}

int bcm_factor_ic2(bcm_matrix upper, bcm_params* p, bcm_factor* out) {
  return guarded([&] {
    need(upper && out, "factor_ic2: null argument");
    need(upper->m.symmetric, "factor_ic2: matrix must be symmetric (upper-stored)");
    auto* h = new bcm_factor_s;
    try {
      build_factor_ic2(upper->m.h, to_slv(p), h->f);
    } catch (...) {
      delete h;
      throw;
    }
    if (p) {
      p->timefct = h->f.stats.seconds;
      p->density = h->f.stats.densu;
      p->density2 = h->f.stats.densr;
    }
    *out = h;
  });
}

int bcm_factor_ilu2(bcm_matrix a, bcm_params* p, bcm_factor* out) {
  return guarded([&] {
    need(a && out, "factor_ilu2: null argument");
    need(!a->m.symmetric, "factor_ilu2: matrix must be stored in full");
    auto* h = new bcm_factor_s;
    try {
      build_factor_ilu2(a->m.h, to_slv(p), h->f);
    } catch (...) {
      delete h;
      throw;
    }
    if (p) {
      p->timefct = h->f.stats.seconds;
      p->density = h->f.stats.densu;
      p->density2 = h->f.stats.densr;
    }
    *out = h;
  });
}

int bcm_factor_destroy(bcm_factor f) {
  return guarded([&] { delete f; });
}

int bcm_factor_info(bcm_factor f, int which, int* nzja, long long* nza, int* levels_fwd, int* levels_bwd) {
  return guarded([&] {
    need(f != nullptr && (which == 0 || which == 1), "factor_info: bad argument");
    const auto& m = factor_part(f->f, which);
    if (nzja) *nzja = m.nzja();
    if (nza) *nza = (long long)m.nza();
    auto levels = [&](bool fwd) {
      std::vector<int> lev(m.nsup, 0);
      int mx = 0;
      if (fwd) {
        for (int i = 0; i < m.nsup; ++i) {
          for (int j = m.ia[i] + 1; j < m.ia[i + 1]; ++j) lev[m.ja[j]] = std::max(lev[m.ja[j]], lev[i] + 1);
          mx = std::max(mx, lev[i]);
        }
      } else {
        for (int i = m.nsup - 1; i >= 0; --i) {
          int l = 0;
          for (int j = m.ia[i] + 1; j < m.ia[i + 1]; ++j) l = std::max(l, lev[m.ja[j]] + 1);
          lev[i] = l;
          mx = std::max(mx, l);
        }
      }
      return m.nsup ? mx + 1 : 0;
    };
    if (levels_fwd) *levels_fwd = levels(true);
    if (levels_bwd) *levels_bwd = levels(false);
  });
}

int bcm_factor_export(bcm_factor f, int which, int* ia, int* ja, double* a) {
  return guarded([&] {
    need(f != nullptr && (which == 0 || which == 1), "factor_export: bad argument");
    const auto& m = factor_part(f->f, which);
    if (ia) std::copy(m.ia.begin(), m.ia.end(), ia);
    if (ja) std::copy(m.ja.begin(), m.ja.end(), ja);
    if (a) std::copy(m.a.begin(), m.a.end(), a);
  });
}

int bcm_device_count(int* count) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    if (count) *count = n;
  });
}

int bcm_shutdown(void) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return;
    BCM_HIP(hipDeviceSynchronize());
    bcm::release_staging();
  });
}

int bcm_dev_alloc(unsigned long long bytes, void** out) {
  return guarded([&] {
    need(out != nullptr, "dev_alloc: out is null");
    need_device();
    *out = nullptr;
    bcm::SetupGuard g;
    BCM_HIP(hipMalloc(out, bytes ? bytes : 1));
  });
}

int bcm_dev_free(void* p) {
  return guarded([&] {
    if (!p) return;
    bcm::SetupGuard g;
    BCM_HIP(hipFree(p));
  });
}

int bcm_copy_to_device(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] {
    need(dst && src, "copy_to_device: null pointer");
    if (!bytes) return;
    bcm::SetupGuard g;
    BCM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  });
}

int bcm_copy_to_host(void* dst, const void* src, unsigned long long bytes) {
  return guarded([&] {
    need(dst && src, "copy_to_host: null pointer");
    if (!bytes) return;
    bcm::SetupGuard g;
    BCM_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  });
}

int bcm_dev_zero(void* p, unsigned long long bytes) {
  return guarded([&] {
    need(p != nullptr, "dev_zero: null pointer");
    if (bytes) {
      // the null stream does not order with the library's non-blocking streams: complete here
      bcm::SetupGuard g;
      BCM_HIP(hipMemset(p, 0, bytes));
      BCM_HIP(hipDeviceSynchronize());
    }
  });
}

int bcm_context_create(bcm_context* out) {
  return guarded([&] {
    need(out != nullptr, "context_create: out is null");
    need_device();
    *out = new bcm_context_s;
  });
}

int bcm_context_destroy(bcm_context c) {
  return guarded([&] { delete c; });
}

int bcm_context_sync(bcm_context c) {
  return guarded([&] {
    need(c != nullptr, "context_sync: null handle");
    BCM_HIP(hipStreamSynchronize(c->c.stream));
    c->c.check_sweep_error();
  });
}

int bcm_context_set_sweep_mode(bcm_context c, int mode) {
  return guarded([&] {
    need(c != nullptr && mode >= 0 && mode <= 2, "context_set_sweep_mode: bad argument");
    BCM_HIP(hipStreamSynchronize(c->c.stream));
    c->c.sweep_mode = mode == 0 ? 0 : 1;
    c->c.sync.level_order = mode == 1;
  });
}

int bcm_context_set_sweep_waves(bcm_context c, int waves_per_cu) {
  return guarded([&] {
    need(c != nullptr && waves_per_cu >= 1 && waves_per_cu <= 64, "context_set_sweep_waves: bad argument");
    BCM_HIP(hipStreamSynchronize(c->c.stream));
    c->c.sync.waves_per_cu = waves_per_cu;
  });
}

int bcm_context_raise_wait_flag(bcm_context c) {
  return guarded([&] {
    need(c != nullptr, "context_raise_wait_flag: null handle");
    bcm::Context& x = c->c;
    x.sync.ensure(0, x.stream);
    const int one = 1;
    bcm::SetupGuard g;
    BCM_HIP(hipMemcpyAsync(x.sync.ctl.as<int>() + 2, &one, sizeof(int), hipMemcpyHostToDevice, x.stream));
    BCM_HIP(hipStreamSynchronize(x.stream));
  });
}

int bcm_context_profile(bcm_context c, int enable) {
  return guarded([&] {
    need(c != nullptr, "context_profile: null handle");
    if (!enable) c->c.prof_flush();
    c->c.prof = enable != 0;
  });
}

int bcm_context_profile_read(bcm_context c, bcm_profile* out, int reset) {
  return guarded([&] {
    need(c && out, "context_profile_read: null argument");
    c->c.prof_flush();
    c->c.prof_collect();
    const auto& sp = c->c.prof_acc[bcm::kProfSpmv];
    const auto& sw = c->c.prof_acc[bcm::kProfSweep];
    out->spmv_ms = sp.ms;
    out->spmv_launches = sp.launches;
    out->spmv_bytes = sp.bytes;
    out->sweep_ms = sw.ms;
    out->sweep_launches = sw.launches;
    out->sweep_bytes = sw.bytes;
    if (reset)
      for (auto& a : c->c.prof_acc) a = bcm::ProfAcc();
  });
}

int bcm_mvm(bcm_context c, bcm_matrix a, const double* x, double* y, int nrhs) {
  return guarded([&] {
    need(c && a && x && y && nrhs > 0, "mvm: bad argument");
    a->m.to_device(false);
    bcm::ProfScope ps(c->c, bcm::kProfSpmv, bcm::spmv_bytes(a->m.d, nrhs));
    bcm::launch_spmv(a->m.d, x, y, nullptr, nrhs, nullptr, nullptr, c->c.rw, c->c.stream);
  });
}

int bcm_mvm_t(bcm_context c, bcm_matrix a, const double* x, double* y, int nrhs) {
  return guarded([&] {
    need(c && a && x && y && nrhs > 0, "mvm_t: bad argument");
    a->m.to_device(true);
    bcm::launch_spmv(a->m.dt, x, y, nullptr, nrhs, nullptr, nullptr, c->c.rw, c->c.stream);
  });
}

int bcm_residual(bcm_context c, bcm_matrix a, const double* b, const double* x, double* r, int nrhs) {
  return guarded([&] {
    need(c && a && b && x && r && nrhs > 0, "residual: bad argument");
    a->m.to_device(false);
    bcm::launch_spmv(a->m.d, x, r, b, nrhs, nullptr, nullptr, c->c.rw, c->c.stream);
  });
}

int bcm_solve_l(bcm_context c, bcm_factor f, int which, const double* x, double* y, int nrhs) {
  return guarded([&] {
    need(c && f && x && y && nrhs > 0 && (which == 0 || which == 1), "solve_l: bad argument");
    need(x != y, "solve_l: x and y must not alias");
    const bool t = which == 1 && f->f.kind == 1;
    const auto& im = f->f.to_device(bcm::kF64, t);
    bcm::launch_sweep(t ? im.fwdt : im.fwd, x, y, bcm::kF64, nrhs, c->c.stream, nullptr, c->c.sweep_sync());
  });
}

int bcm_solve_u(bcm_context c, bcm_factor f, int which, const double* x, double* y, int nrhs) {
  return guarded([&] {
    need(c && f && x && y && nrhs > 0 && (which == 0 || which == 1), "solve_u: bad argument");
    need(x != y, "solve_u: x and y must not alias");
    const bool t = which == 0 && f->f.kind == 1;
    const auto& im = f->f.to_device(bcm::kF64, t);
    bcm::launch_sweep(t ? im.bwdt : im.bwd, x, y, bcm::kF64, nrhs, c->c.stream, nullptr, c->c.sweep_sync());
  });
}

int bcm_precond(bcm_context c, bcm_factor f, const double* r, double* z, int nrhs, int precision, int transpose) {
  return guarded([&] {
    need(c && f && r && z && nrhs > 0, "precond: bad argument");
    need(precision >= 0 && precision <= 2, "precond: bad precision");
    bcm::precond_apply(f->f, c->c, r, z, precision, transpose != 0, nrhs);
  });
}

int bcm_precond_noalloc(bcm_context c, bcm_factor f, const double* r, double* z, int nrhs, int precision,
                        int transpose) {
  return guarded([&] {
    need(c && f && r && z && nrhs > 0, "precond_noalloc: bad argument");
    need(r != z, "precond_noalloc: r and z must not alias");
    bcm::precond_noalloc(f->f, c->c, r, z, precision, transpose != 0, nrhs);
  });
}

int bcm_dot(bcm_context c, int n, const double* x, const double* y, double* result) {
  return guarded([&] {
    need(c && x && y && result && n >= 0, "dot: bad argument");
    bcm::KState* ks = c->c.dstate();
    const double* xs[1] = {x};
    const double* ys[1] = {y};
    bcm::launch_multidot(n, 1, xs, ys, &ks->sc[31], c->c.rw, c->c.stream);
    c->c.pull_state();
    *result = c->c.host_state->sc[31];
  });
}

int bcm_solve(bcm_context c, bcm_matrix a, bcm_factor f, const bcm_solve_opts* o, const double* b, double* x,
              bcm_stats* st) {
  return guarded([&] {
    need(c && a && f && b && x, "solve: bad argument");
    bcm::SolveOptions opt = to_opts(o);
    bcm::SolveStats s;
    bcm::solve(a->m, f->f, c->c, opt, b, x, s);
    fill_stats(s, st);
    if (s.breakdown) throw bcm::Error(bcm::kErrBreakdown, "Krylov breakdown");
  });
}

int bcm_solve_block(bcm_context c, bcm_matrix a, bcm_factor f, const bcm_solve_opts* o, int nrhs, const double* b,
                    double* x, bcm_stats* st, double* resi, double* hist) {
  return guarded([&] {
    need(c && a && f && b && x, "solve_block: bad argument");
    need(nrhs >= 1 && nrhs <= 8, "solve_block: 1..8 right-hand sides");
    bcm::SolveOptions opt = to_opts(o);
    const int v = bcm::blz_variant(opt.kind);
    if (v < 0) opt.kind = bcm::kBlockGmres;
    bcm::SolveStats s;
    auto t0 = std::chrono::steady_clock::now();
    if (v >= 0) bcm::solve_block_lanczos(a->m, f->f, c->c, opt, v, nrhs, b, x, s, resi);
    else bcm::solve_block_gmres(a->m, f->f, c->c, opt, nrhs, b, x, s, resi, hist);
    BCM_HIP(hipStreamSynchronize(c->c.stream));
    c->c.check_sweep_error();
    s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    fill_stats(s, st);
    if (s.breakdown) throw bcm::Error(bcm::kErrBreakdown, "Krylov breakdown");
  });
}

int bcm_batch_create(int ncell, const bcm_matrix* mats, const bcm_factor* facts, int prec_mask, bcm_batch* out) {
  return guarded([&] {
    need(ncell > 0 && mats && facts && out, "batch_create: bad argument");
    need((prec_mask & 7) != 0, "batch_create: prec_mask selects no precision");
    need_device();
    std::vector<const bcm::Matrix*> m(ncell);
    std::vector<const bcm::Factor*> f(ncell);
    for (int i = 0; i < ncell; ++i) {
      need(mats[i] && facts[i], "batch_create: null handle");
      m[i] = &mats[i]->m;
      f[i] = &facts[i]->f;
    }
    std::unique_ptr<bcm_batch_s> b(new bcm_batch_s());
    bcm::batch_build(b->b, m, f, prec_mask & 7);
    *out = b.release();
  });
}

int bcm_batch_destroy(bcm_batch b) {
  delete b;
  return BCM_OK;
}

int bcm_batch_layout(bcm_batch b, int* ncell, long long* n, long long* elem0) {
  return guarded([&] {
    need(b != nullptr, "batch_layout: bad argument");
    if (ncell) *ncell = b->b.ncell;
    if (n) *n = b->b.n;
    if (elem0)
      for (int i = 0; i <= b->b.ncell; ++i) elem0[i] = b->b.elem0[i];
  });
}

int bcm_batch_solve(bcm_context c, bcm_batch b, const bcm_solve_opts* o, int nrhs, const double* rhs, double* x,
                    int fallback, bcm_batch_stats* st, int* iters, double* relres) {
  return guarded([&] {
    need(c && b && rhs && x, "batch_solve: bad argument");
    bcm::SolveOptions opt = to_opts(o);
    if (opt.kind != bcm::kPcg) opt.kind = bcm::kBlockPcg;
    bcm::BatchStats s;
    need(fallback >= 0 && fallback <= 3, "batch_solve: start flags are bits 0 (fallback) and 1 (Galerkin)");
    bcm::batch_solve(b->b, c->c, opt, nrhs, rhs, x, (fallback & 1) != 0, s, iters, relres, (fallback & 2) != 0);
    if (st) {
      st->pairs = s.pairs;
      st->converged = s.converged;
      st->breakdowns = s.breakdowns;
      st->iters_max = s.iters_max;
      st->restarts = s.restarts;
      st->iters_sum = s.iters_sum;
      st->relres_max = s.relres_max;
      st->seconds = s.seconds;
      st->recovered = s.recovered;
    }
  });
}

static void tune_out(const bcm::BatchTune& t, bcm_batch_tune* out) {
  out->sweep_mode = t.sweep_mode;
  out->waves_per_cu = t.waves_per_cu;
  out->level_order = t.level_order;
  out->graph = t.graph;
  out->check_every = t.check_every;
  out->ms_per_iter = t.ms_per_iter;
}

int bcm_batch_autotune(bcm_context c, bcm_batch b, const bcm_solve_opts* o, int nrhs, const double* rhs,
                       int allow_graph, bcm_batch_tune* out) {
  return guarded([&] {
    need(c && b && rhs, "batch_autotune: bad argument");
    bcm::SolveOptions opt = to_opts(o);
    opt.kind = bcm::kBlockPcg;
    bcm::BatchTune t = bcm::batch_tune(b->b, c->c, opt, nrhs, rhs, allow_graph != 0);
    if (out) tune_out(t, out);
  });
}

int bcm_batch_get_tune(bcm_batch b, bcm_batch_tune* t) {
  return guarded([&] {
    need(b && t, "batch_get_tune: bad argument");
    tune_out(b->b.tune, t);
  });
}

int bcm_batch_set_tune(bcm_batch b, const bcm_batch_tune* t) {
  return guarded([&] {
    need(b && t, "batch_set_tune: bad argument");
    need(t->sweep_mode == 0 || t->sweep_mode == 1, "batch_set_tune: sweep_mode is 0 or 1");
    need(t->waves_per_cu >= 1 && t->waves_per_cu <= 32, "batch_set_tune: waves_per_cu in 1..32");
    need(t->check_every >= 1 && t->check_every <= 1024, "batch_set_tune: check_every in 1..1024");
    b->b.tune.sweep_mode = t->sweep_mode;
    b->b.tune.waves_per_cu = t->waves_per_cu;
    b->b.tune.level_order = t->level_order ? 1 : 0;
    b->b.tune.graph = t->graph ? 1 : 0;
    b->b.tune.check_every = t->check_every;
    b->b.drop_graph();
  });
}

int bcm_mlp_create(int nin, int nhidden, int nout, unsigned long long seed, bcm_mlp* out) {
  return guarded([&] {
    need(out != nullptr, "mlp_create: bad argument");
    need_device();
    std::unique_ptr<bcm_mlp_s> m(new bcm_mlp_s());
    m->m.reset(new bcm::Mlp(nin, nhidden, nout, seed));
    *out = m.release();
  });
}

int bcm_mlp_destroy(bcm_mlp m) {
  delete m;
  return BCM_OK;
}

int bcm_mlp_nparam(bcm_mlp m, long long* nparam) {
  return guarded([&] {
    need(m && nparam, "mlp_nparam: bad argument");
    *nparam = m->m->nparam;
  });
}

int bcm_mlp_set_params(bcm_mlp m, const float* host) {
  return guarded([&] {
    need(m && host, "mlp_set_params: bad argument");
    m->m->set_params(host);
  });
}

int bcm_mlp_get_params(bcm_mlp m, float* host) {
  return guarded([&] {
    need(m && host, "mlp_get_params: bad argument");
    BCM_HIP(hipDeviceSynchronize());
    BCM_HIP(hipMemcpy(host, m->m->params.p, m->m->nparam * 4, hipMemcpyDeviceToHost));
  });
}

int bcm_mlp_get_grads(bcm_mlp m, float* host) {
  return guarded([&] {
    need(m && host, "mlp_get_grads: bad argument");
    BCM_HIP(hipDeviceSynchronize());
    BCM_HIP(hipMemcpy(host, m->m->grads.p, m->m->nparam * 4, hipMemcpyDeviceToHost));
  });
}

int bcm_mlp_forward(bcm_context c, bcm_mlp m, int batch, const float* x, float* y) {
  return guarded([&] {
    need(c && m && x && y && batch > 0, "mlp_forward: bad argument");
    m->m->forward(c->c.stream, batch, x, y);
  });
}

int bcm_mlp_backward(bcm_context c, bcm_mlp m, int batch, const float* x, const float* t, double* loss) {
  return guarded([&] {
    need(c && m && x && t && batch > 0, "mlp_backward: bad argument");
    double* dl = c->c.vec(60, 16);
    m->m->backward(c->c.stream, batch, x, t, dl);
    BCM_HIP(hipStreamSynchronize(c->c.stream));
    if (loss) BCM_HIP(hipMemcpy(loss, dl, 8, hipMemcpyDeviceToHost));
  });
}

int bcm_mlp_adam_step(bcm_context c, bcm_mlp m, float lr) {
  return guarded([&] {
    need(c && m, "mlp_adam_step: bad argument");
    m->m->adam(c->c.stream, lr);
    BCM_HIP(hipStreamSynchronize(c->c.stream));
  });
}

int bcm_guess_create(long long n, int nrhs, int nparam, int nmodes, int nhidden, int capacity,
                     unsigned long long seed, bcm_guess* out) {
  return guarded([&] {
    need(out != nullptr, "guess_create: bad argument");
    need_device();
    std::unique_ptr<bcm_guess_s> g(new bcm_guess_s());
    g->g.reset(new bcm::Guess(n, nrhs, nparam, nmodes, nhidden, capacity, seed));
    *out = g.release();
  });
}

int bcm_guess_destroy(bcm_guess g) {
  delete g;
  return BCM_OK;
}

int bcm_guess_add(bcm_context c, bcm_guess g, int count, const double* params, const double* x) {
  return guarded([&] {
    need(c && g && params && x, "guess_add: bad argument");
    g->g->add(c->c.stream, count, params, x);
  });
}

int bcm_guess_fit(bcm_context c, bcm_guess g, int steps, double lr, bcm_guess_report* rep) {
  return guarded([&] {
    need(c && g && steps >= 0 && lr > 0, "guess_fit: bad argument");
    bcm::GuessReport r = g->g->fit(c->c.stream, steps, lr);
    if (rep) {
      rep->snapshots = r.snapshots;
      rep->modes = r.modes;
      rep->steps = r.steps;
      rep->loss_first = r.loss_first;
      rep->loss_final = r.loss_final;
      rep->pod_energy = r.pod_energy;
      rep->seconds_pod = r.seconds_pod;
      rep->seconds_train = r.seconds_train;
    }
  });
}

int bcm_guess_predict(bcm_context c, bcm_guess g, int count, const double* params, double* x0) {
  return guarded([&] {
    need(c && g && params && x0, "guess_predict: bad argument");
    g->g->predict(c->c.stream, count, params, x0);
  });
}

int* bcm_alloc_ints(int nint) { return new (std::nothrow) int[nint > 0 ? nint : 1]; }

}  // extern "C"
