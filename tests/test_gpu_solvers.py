"""GPU Krylov drivers through the C ABI against the CPU oracle.

Parity bar: relative solution difference <= 1e-8 against the oracle on the
same seeded inputs, FP64 true relative residual <= the requested tolerance,
and the oracle's iteration count (PCG: exact; BiCGStab/FGMRES: +-1, the
recurrences round differently).
"""

import ctypes as C

import numpy as np
import pytest

from conftest import rel
from helpers import heat, seepage

pytestmark = pytest.mark.gpu


def dev(B, x):
    return B.DeviceArray.from_host(np.asfortranarray(x))


def solve(B, ctx, mo, f, b, **opts):
    db = dev(B, b)
    dx = B.DeviceArray(len(b))
    dx.zero()
    st = ctx.solve(mo, f, db, dx, B.default_opts(**opts))
    return dx.host(), st


@pytest.mark.parametrize("ittype,eps", [(0, 1e-10), (3, 1e-10), (1, 1e-6), (2, 1e-9), (2, -1e-7)])
def test_pcg_matches_oracle(B, O, ctx, ittype, eps):
    """Serial Cg (cg.cpp:31-349) under each stopping rule, ittype 2 being the
    reference default: max_i |r_i| <= |eps| (cg.cpp:241 over MaxProdSVect,
    svector.cpp:376-385, with the unit norm vector of solvers.cpp:5554)."""
    h = heat(B, O, 16, 14, 9, seed=11)
    print(f"pcg ittype={ittype} eps={eps} n={h['A'].n}", flush=True)
    x, st = solve(B, ctx, h["mo"], h["f"], h["b"], kind=B.PCG, eps=eps, ittype=ittype, check_every=3)
    xo, r = O.cg(h["A"], h["U"], h["b"], eps=eps, ittype=ittype)
    assert st.converged == 1 and st.breakdown == 0
    assert st.iterations == r.updates
    assert rel(x, xo) < 1e-8 * max(1.0, abs(eps) / 1e-10)
    if ittype == 3:
        assert st.relres_true <= eps
    if ittype == 2:
        # errmin of the exit check: the max-norm of the GPU's residual, which
        # tracks the oracle's to the rounding of the parallel dots
        r_true = h["b"] - h["mo"].dense() @ x
        assert st.errmin <= abs(eps) and abs(st.errmin - r.resfin) <= 1e-3 * abs(eps)
        assert np.max(np.abs(r_true)) <= 10 * abs(eps)


def test_bicgstab_matches_oracle(B, O, ctx):
    s = seepage(B, O, 6, 5, 4, 24, seed=12)
    x, st = solve(B, ctx, s["mo"], s["f"], s["b"], kind=B.BICGSTAB, eps=1e-10)
    xo, r = O.bicgstab(s["A"], s["L"], s["U"], s["b"], eps=1e-10)
    assert st.converged == 1 and st.relres_true <= 1e-10
    assert abs(st.iterations - r.nitperf) <= 1
    assert rel(x, xo) < 1e-8


def test_pcg_warm_start_and_zero_rhs(B, O, ctx):
    h = heat(B, O, 10, 10, 9, seed=13)
    xd = np.linalg.solve(h["mo"].dense(), h["b"])
    db, dx = dev(B, h["b"]), dev(B, xd * (1 + 1e-6))
    st = ctx.solve(h["mo"], h["f"], db, dx, B.default_opts(kind=B.PCG, eps=1e-10))
    assert st.converged and st.iterations < 15 and st.relres_true <= 1e-10
    z = np.zeros(h["A"].n)
    x, st = solve(B, ctx, h["mo"], h["f"], z, kind=B.PCG, eps=1e-10)
    assert np.all(x == 0) and st.iterations == 0


def _callbacks(B, m, rhs, store):
    sprnd, ia, ja, a = m.arrays()
    info = m.info()
    nsup = info["nsup"]
    rowstart = np.zeros(nsup + 1, np.int64)
    for i in range(nsup):
        bi = sprnd[i + 1] - sprnd[i]
        rowstart[i + 1] = rowstart[i] + sum(bi * (sprnd[c + 1] - sprnd[c]) for c in ja[ia[i]:ia[i + 1]])
    perm = {i: np.random.default_rng(i).permutation(ia[i + 1] - ia[i]) for i in range(nsup)}

    def fnsup(obj, p):
        p[0] = nsup

    def fsprnds(obj, p):
        for i in range(nsup + 1):
            p[i] = int(sprnd[i])

    def fiaja(obj, pia, pja):
        ib = B.lib().bcm_alloc_ints(nsup + 1)
        jb = B.lib().bcm_alloc_ints(len(ja))
        for i in range(nsup + 1):
            ib[i] = int(ia[i])
        for i in range(nsup):   # callback order differs from sorted order: the library must sort
            for k, src in enumerate(perm[i]):
                jb[ia[i] + k] = int(ja[ia[i] + src])
        pia[0] = ib
        pja[0] = jb

    def fblock(obj, i, out):
        bi = sprnd[i + 1] - sprnd[i]
        off = 0
        offs = []
        for c in ja[ia[i]:ia[i + 1]]:
            offs.append(off)
            off += bi * (sprnd[c + 1] - sprnd[c])
        pos = 0
        for src in perm[i]:
            c = ja[ia[i] + src]
            sz = bi * (sprnd[c + 1] - sprnd[c])
            blk = a[rowstart[i] + offs[src]: rowstart[i] + offs[src] + sz]
            for t in range(sz):
                out[pos + t] = blk[t]
            pos += sz

    def frhs(obj, i, out):
        for t in range(sprnd[i], sprnd[i + 1]):
            out[t - sprnd[i]] = rhs[t]

    def fini(obj, i, out):
        for t in range(sprnd[i + 1] - sprnd[i]):
            out[t] = 0.0

    def fstore(obj, i, out):
        for t in range(sprnd[i], sprnd[i + 1]):
            store[t] = out[t - sprnd[i]]

    sig = B.SIGNATURES["bcm_abstract_symm_solver"][1]
    return [sig[2](fnsup), sig[3](fsprnds), sig[4](fiaja), sig[5](fblock), sig[6](frhs), sig[7](fini),
            sig[8](fstore)]


@pytest.mark.parametrize("symmetric,ittype", [(True, 3), (True, 2), (False, 1)],
                         ids=["symm-pcg", "symm-pcg-maxnorm", "lanczos"])
def test_dropin_abstract_solvers(B, O, symmetric, ittype):
    if symmetric:
        m, rhs = B.gen_heat2d(9, 7, 9, 31)
    else:
        m, rhs = B.gen_seepage3d(4, 4, 3, 6, 0.3, 32)
    x = np.zeros(m.n)
    cbs = _callbacks(B, m, rhs, x)
    eps = 1e-11 if ittype == 2 else 1e-10   # ittype 2: absolute max-norm of r
    p = B.default_params(ittype=ittype, sttype=1, eps=eps, niter=500)
    fn = B.lib().bcm_abstract_symm_solver if symmetric else B.lib().bcm_abstract_solver
    code = fn(None, C.byref(p), *cbs)
    assert code == 0, B.lib().bcm_last_error()
    xd = np.linalg.solve(m.dense(), rhs)
    assert rel(x, xd) < 1e-8
    assert p.nitperf > 0 and p.timefct > 0 and p.density >= 1.0
    if symmetric:  # nitperf as the reference's Cg reports it: its loop index at exit (cg.cpp:209)
        order = m.rcm_order()
        mo = m.reorder(order)
        f = B.Factor.ic2(mo, B.default_params(ittype=ittype, sttype=1, eps=1e-10, niter=500))
        A = O.Block(*mo.arrays())
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
        _, r = O.cg(A, U, mo.reorder_vector(order, rhs), eps=eps, ittype=ittype)
        assert p.nitperf == r.nitperf, (p.nitperf, r.nitperf)
        if ittype == 2:   # resfin = errmin (cg.cpp:210)
            assert p.resfin <= eps and abs(p.resfin - r.resfin) <= 1e-3 * eps


@pytest.mark.parametrize("ittype,sttype,eps", [(0, 1, 1e-12), (2, 2, 1e-11), (1, 0, 1e-11)])
def test_lanczos_matches_oracle(B, O, ctx, ittype, sttype, eps):
    """Two-sided Lanczos (lanczos.cpp:34-371): same iterations as the oracle's
    restatement (+-1: the dots are parallel sums), the same good iterate to
    1e-8 and an FP64 true relative residual within 1e-10. The tolerances are
    set where the recurrence is converged to rounding (1e-12 relative to
    ||r0||, 1e-11 absolute with ||b|| ~ 44), so that one iteration more or
    less moves the iterate by far less than the 1e-8 bar."""
    s = seepage(B, O, 5, 4, 4, 24, seed=31)
    print(f"lanczos ittype={ittype} eps={eps} n={s['A'].n}", flush=True)
    x, st = solve(B, ctx, s["mo"], s["f"], s["b"], kind=B.LANCZOS, eps=eps, ittype=ittype, check_every=2)
    xo, r = O.lanczos(s["A"], s["L"], s["U"], s["b"], eps=eps, sttype=sttype)
    assert st.converged == 1 and st.breakdown == 0
    assert abs(st.iterations - r.nitperf) <= 1, (st.iterations, r.nitperf)
    assert rel(x, xo) < 1e-8
    assert st.relres_true <= 1e-10
    r_true = s["b"] - s["mo"].dense() @ x
    assert np.linalg.norm(r_true) <= 1e-10 * np.linalg.norm(s["b"])


@pytest.mark.parametrize("ittype,sttype", [(0, 1), (1, 0)])
def test_cg_par_matches_oracle(B, O, ctx, ittype, sttype):
    """Parallel-form CG (cg.cpp:355-771) against the oracle's restatement (with
    the last solution step applied at a converged exit)."""
    h = heat(B, O, 16, 14, 9, seed=43)
    eps = 1e-10 if sttype == 1 else 1e-8
    x, st = solve(B, ctx, h["mo"], h["f"], h["b"], kind=B.CGPAR, eps=eps, ittype=ittype, check_every=4)
    xo, r = O.cg_par(h["A"], h["U"], h["b"], eps=eps, sttype=sttype, apply_last=True)
    assert st.converged == 1 and st.breakdown == 0
    assert abs(st.iterations - r.updates) <= 1, (st.iterations, r.updates)
    assert rel(x, xo) < 1e-8 * max(1.0, eps / 1e-10)


@pytest.mark.parametrize("ittype,skew", [(1, 0.0), (2, 0.3), (0, 0.3)], ids=["cg-par", "lanczos", "block-gmres"])
def test_dropin_abstract_par_solver(B, O, ittype, skew):
    """AbstractParSolver on one rank: the kind of Solver(tree)
    (solvers.cpp:1865-2041) by ittype (1 parallel Cg on a symmetric system in
    full storage, 2 Lanczos, otherwise block GMRES); the stored solution
    solves the system."""
    m, rhs = B.gen_seepage3d(4, 4, 3, 6, skew, 33)
    x = np.zeros(m.n)
    cbs = _callbacks(B, m, rhs, x)
    p = B.default_params(ittype=ittype, sttype=1, eps=1e-11, niter=400)
    code = B.lib().bcm_abstract_par_solver(None, None, C.byref(p), *cbs)
    assert code == 0, B.lib().bcm_last_error()
    Ad = m.dense()
    xd = np.linalg.solve(Ad, rhs)
    assert rel(x, xd) < 1e-8, ittype
    assert np.linalg.norm(rhs - Ad @ x) <= 1e-9 * np.linalg.norm(rhs)
    assert p.nitperf > 0 and p.timefct > 0


@pytest.mark.parametrize("ittype,skew", [(1, 0.0), (0, 0.3)], ids=["cg-par", "block-gmres"])
def test_dropin_par_solver_on_a_dissected_system(B, O, ittype, skew):
    """720 block rows: the parallel solver's nested dissection splits the
    pattern (several tree levels below the 64-row leaves); the solve on the
    dissected, ILU2-factored system returns the solution in the caller's
    ordering."""
    m, rhs = B.gen_seepage3d(10, 9, 8, 6, skew, 34)
    order, blks, _ = m.part_binary_tree_nd(8)
    assert len(blks) > 8  # the tree did split
    x = np.zeros(m.n)
    cbs = _callbacks(B, m, rhs, x)
    p = B.default_params(ittype=ittype, sttype=1, eps=1e-11, niter=400)
    print(f"par solver ittype={ittype} n={m.n}", flush=True)
    code = B.lib().bcm_abstract_par_solver(None, None, C.byref(p), *cbs)
    assert code == 0, B.lib().bcm_last_error()
    Ad = m.dense()
    xd = np.linalg.solve(Ad, rhs)
    assert rel(x, xd) < 1e-8, ittype
    assert np.linalg.norm(rhs - Ad @ x) <= 1e-9 * np.linalg.norm(rhs)
