"""Complex solve phase on the GPU through its real equivalent (bcm_cmatrix_create,
bcm_cfactor_create): the device products and sweeps against the reference's
own complex MvmA / SolveL / SolveU (mvmc.cpp:160, 1723, 2506, compiled in
oracle/_ref), and complete preconditioned solves against a dense complex solve.

Parity bar: operators within 1e-13 of the reference's complex arithmetic
(rounding-level: the real form sums real and imaginary products in separate
column runs); solutions within 1e-8 of the dense solution and FP64 true
relative residual within the requested tolerance. The factors are the
reference's own complex ILU2 (CombineLU + CSMatrixCS::Ilu2, ilu2cs.cpp).
"""

import ctypes as C

import numpy as np
import pytest

from test_complex import complex_system, dense_complex

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csys(B, O):
    sp, ia, ja, ac = complex_system(B, 8, 6, 5, 12, seed=21)  # complex blocks of 12: real blocks of 24
    L, U = O.ref_cilu2(sp, ia, ja, ac)
    return dict(sp=sp, ia=ia, ja=ja, ac=ac, L=L, U=U, mc=B.Matrix.from_complex(sp, ia, ja, ac),
                f=B.Factor.from_complex(sp, L, U))


def dev(B, x):
    return B.DeviceArray.from_host(np.asfortranarray(x))


@pytest.mark.parametrize("nrhs", [1, 2, 3])
def test_complex_operators_match_reference(B, O, ctx, csys, nrhs):
    s = csys
    sp = s["sp"]
    n = sp[-1]
    print(f"complex operators n={n} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(7 + nrhs)
    z = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    dx = dev(B, B.complex_to_real(sp, z))
    dy = B.DeviceArray((2 * n, nrhs))
    ctx.mvm(s["mc"], dx, dy, nrhs)
    ctx.sync()
    y = B.real_to_complex(sp, dy.host())
    yr = O.ref_cmvma(sp, s["ia"], s["ja"], s["ac"], z, nrhs)
    assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max()
    for mode in (1, 0):
        ctx.set_sweep_mode(mode)
        for which, fn, part in ((0, ctx.solve_l, s["L"]), (1, ctx.solve_u, s["U"])):
            fn(s["f"], which, dx, dy, nrhs)
            ctx.sync()
            y = B.real_to_complex(sp, dy.host())
            yr = O.ref_csolve(which, sp, *part, z, nrhs)
            assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max(), (mode, which)
    ctx.set_sweep_mode(1)


@pytest.mark.parametrize("bc", [3, 4, 6, 8, 16])
def test_complex_block_kernels_every_block_size(B, O, ctx, bc):
    """The complex-block SpMV and sweeps (complex blocks of 3..16 read as
    (re, im) pairs) at 1, 2, 6 and 7 columns (column chunks of 1, 2, 3 and
    6): products against the reference's complex MvmA, y = b - A x (the
    residual form the solvers launch) through bcm_residual, and both sweep
    schedules against the reference's complex SolveL / SolveU on the
    reference's own complex ILU2 factors."""
    sp, ia, ja, ac = complex_system(B, 4, 3, 3, bc, seed=40 + bc)
    mc = B.Matrix.from_complex(sp, ia, ja, ac)
    L, U = O.ref_cilu2(sp, ia, ja, ac)
    f = B.Factor.from_complex(sp, L, U)
    n = sp[-1]
    for nrhs in (1, 2, 6, 7):
        print(f"complex kernels bc={bc} n={n} nrhs={nrhs}", flush=True)
        rng = np.random.default_rng(bc * 10 + nrhs)
        z = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        dx = dev(B, B.complex_to_real(sp, z))
        dy = B.DeviceArray((2 * n, nrhs))
        ctx.mvm(mc, dx, dy, nrhs)
        ctx.sync()
        y = B.real_to_complex(sp, dy.host())
        yr = O.ref_cmvma(sp, ia, ja, ac, z, nrhs)
        assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max(), nrhs
        w = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
        ctx.residual(mc, dev(B, B.complex_to_real(sp, w)), dx, dy, nrhs)
        ctx.sync()
        r = B.real_to_complex(sp, dy.host())
        assert np.abs(r - (w - yr)).max() <= 1e-13 * np.abs(yr).max(), nrhs
        for mode in (1, 2, 0):  # level-ordered sync-free, natural-order sync-free, one launch per level
            ctx.set_sweep_mode(mode)
            for which, fn, part in ((0, ctx.solve_l, L), (1, ctx.solve_u, U)):
                fn(f, which, dx, dy, nrhs)
                ctx.sync()
                y = B.real_to_complex(sp, dy.host())
                yr = O.ref_csolve(which, sp, *part, z, nrhs)
                assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max(), (nrhs, mode, which)
        ctx.set_sweep_mode(1)
        # the allocation-free preconditioner shares the complex sweeps: bit for bit
        dp, dq = B.DeviceArray((2 * n, nrhs)), B.DeviceArray((2 * n, nrhs))
        ctx.precond(f, dx, dp, nrhs)
        ctx.precond_noalloc(f, dx, dq, nrhs)
        ctx.sync()
        assert np.array_equal(dp.host(), dq.host()), nrhs


@pytest.mark.parametrize("kind", ["BICGSTAB", "FGMRES", "BLOCKGMRES"])
def test_complex_solve(B, O, ctx, csys, kind):
    """A preconditioned solve of the complex system (ILU2 from the reference)
    by a Krylov driver on its real equivalent: the complex solution of the
    dense system, to the requested FP64 true residual."""
    s = csys
    sp = s["sp"]
    n = sp[-1]
    rng = np.random.default_rng(11)
    bz = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xd = np.linalg.solve(dense_complex(sp, s["ia"], s["ja"], s["ac"]), bz)
    print(f"complex solve {kind} n={n}", flush=True)
    db = dev(B, B.complex_to_real(sp, bz))
    dx = B.DeviceArray(2 * n)
    dx.zero()
    eps = 1e-11
    if kind == "BLOCKGMRES":
        st, resi, _ = ctx.solve_block(s["mc"], s["f"], db, dx, 1,
                                      B.default_opts(kind=B.BLOCKGMRES, eps=eps, niter=200, ittype=0))
        assert st.converged == 1
    else:
        st = ctx.solve(s["mc"], s["f"], db, dx, B.default_opts(kind=getattr(B, kind), eps=eps, ittype=3, niter=500))
        assert st.converged == 1 and st.relres_true <= eps
    x = B.real_to_complex(sp, dx.host())
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)


def _ccallbacks(B, O, sp, ia, ja, ac, bz, x0, store, reference):
    """Callbacks of the complex AbstractSolver (CNSUP ... CSTORESOLC) over a
    complex block matrix, each row's blocks handed out in a permuted order
    (the solver must sort them); complex data as interleaved doubles. With
    reference, the types and index allocator of the reference build."""
    nsup = len(sp) - 1
    av = np.ascontiguousarray(ac, np.complex128)
    bv = np.ascontiguousarray(bz, np.complex128)
    xv = np.ascontiguousarray(x0, np.complex128)
    start = np.zeros(len(ja) + 1, np.int64)  # complex offset of every block in ja order
    for i in range(nsup):
        for j in range(ia[i], ia[i + 1]):
            start[j + 1] = start[j] + (sp[i + 1] - sp[i]) * (sp[ja[j] + 1] - sp[ja[j]])
    perm = [np.random.default_rng(100 + i).permutation(ia[i + 1] - ia[i]) for i in range(nsup)]
    alloc = O.ref_alloc_ints if reference else B.lib().bcm_alloc_ints

    def fnsup(obj, p):
        p[0] = nsup

    def fsprnds(obj, p):
        for i in range(nsup + 1):
            p[i] = int(sp[i])

    def fiaja(obj, pia, pja):
        ib, jb = alloc(nsup + 1), alloc(len(ja))
        for i in range(nsup + 1):
            ib[i] = int(ia[i])
        for i in range(nsup):
            for k, src in enumerate(perm[i]):
                jb[ia[i] + k] = int(ja[ia[i] + src])
        pia[0] = ib
        pja[0] = jb

    def fblock(obj, i, out):
        pos = 0
        for src in perm[i]:
            j = ia[i] + src
            blk = av[start[j]:start[j + 1]]
            C.memmove(C.addressof(out.contents) + 16 * pos, blk.ctypes.data, blk.nbytes)
            pos += len(blk)

    def part(v):
        def f(obj, i, out):
            seg = v[sp[i]:sp[i + 1]]
            C.memmove(C.addressof(out.contents), seg.ctypes.data, seg.nbytes)
        return f

    def fstore(obj, i, out):
        seg = np.ctypeslib.as_array(out, shape=(2 * (sp[i + 1] - sp[i]),))
        store[sp[i]:sp[i + 1]] = seg[0::2] + 1j * seg[1::2]

    fns = [fnsup, fsprnds, fiaja, fblock, part(bv), part(xv), fstore]
    if reference:
        return fns
    t = B.SIGNATURES["bcm_abstract_solver_c"][1]
    return [t[2](fnsup), t[3](fsprnds), t[4](fiaja), t[5](fblock), t[6](part(bv)), t[7](part(xv)), t[8](fstore)]


def _drop_in(B, O, sys_, bz, x0, reference=False, **prm):
    sp, ia, ja, ac = sys_
    x = np.full(sp[-1], np.nan + 0j)
    cbs = _ccallbacks(B, O, sp, ia, ja, ac, bz, x0, x, reference)
    if reference:
        O.ref_abstract_solver_c(*cbs, memory=10.0, **prm)
        return x, None
    p = B.default_params(**prm)
    code = B.lib().bcm_abstract_solver_c(None, C.byref(p), *cbs)
    assert code == 0, B.lib().bcm_last_error()
    return x, p


@pytest.fixture(scope="module", params=["b12", "variable"])
def dsys(request, B):
    if request.param == "variable":
        from test_complex import _variable_system
        return _variable_system()
    return complex_system(B, 6, 5, 4, 12, seed=23)


@pytest.mark.parametrize("ordtype,steps", [(1, 4), (1, 9), (0, 6)])
def test_dropin_complex_fixed_steps_equal_reference(B, O, dsys, ordtype, steps):
    """bcm_abstract_solver_c against the reference's own complex
    AbstractSolver (solvers.cpp:7103) on the same callbacks, with eps = 0 so
    that both run exactly `steps` block GMRES steps: the iterates agree to
    rounding. This pins the assembly, the ordering, the bit-exact complex
    ILU2 and the real block GMRES over [b, i b] against the reference's
    complex block GMRES (gmres.cpp:1044) step for step."""
    sp = dsys[0]
    n = sp[-1]
    rng = np.random.default_rng(5 + steps)
    bz = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = 0.1 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    print(f"complex drop-in n={n} ordtype={ordtype} steps={steps}", flush=True)
    x, p = _drop_in(B, O, dsys, bz, x0, ordtype=ordtype, niter=steps, eps=0.0)
    xr, _ = _drop_in(B, O, dsys, bz, x0, reference=True, ordtype=ordtype, niter=steps, eps=0.0)
    assert p.nitperf == steps
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr), np.linalg.norm(x - xr) / np.linalg.norm(xr)


def test_dropin_complex_converges_with_reference(B, O, dsys):
    """Converged to eps = 1e-11 relative to ||r0||: both solutions within 1e-8
    of the dense complex solution, and the FP64 true relative residual
    reported in resfin."""
    sp, ia, ja, ac = dsys
    n = sp[-1]
    rng = np.random.default_rng(17)
    bz = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x0 = np.zeros(n, complex)
    xd = np.linalg.solve(dense_complex(sp, ia, ja, ac), bz)
    x, p = _drop_in(B, O, dsys, bz, x0, niter=300, eps=1e-11)
    xr, _ = _drop_in(B, O, dsys, bz, x0, reference=True, niter=300, eps=1e-11)
    print(f"complex drop-in converged in {p.nitperf} steps, resfin {p.resfin:.2e}", flush=True)
    assert 0 < p.nitperf < 300 and p.resfin <= 1e-10
    assert np.linalg.norm(x - xd) <= 1e-8 * np.linalg.norm(xd)
    assert np.linalg.norm(xr - xd) <= 1e-8 * np.linalg.norm(xd)
    assert p.density > 0 and p.timefct > 0 and p.timeitr > 0


def test_dropin_complex_zero_rhs_and_bad_ordering(B, O, dsys):
    """A zero right-hand side stores the initial guess unchanged: the
    reference's solver skips zero columns (solvers.cpp:2725-2790), though its
    closing diagnostic ZGESVD of the then empty system stops the process
    (LWORK = 0, parameter 13, through XERBLA), which is not run here. METIS
    orderings (ordtype > 1) are rejected."""
    sp = dsys[0]
    n = sp[-1]
    x0 = np.random.default_rng(3).standard_normal(n) + 0.5j
    x, p = _drop_in(B, O, dsys, np.zeros(n, complex), x0)
    assert np.array_equal(x, x0) and p.nitperf == 0
    cbs = _ccallbacks(B, O, *dsys, np.ones(n, complex), x0, np.zeros(n, complex), False)
    p = B.default_params(ordtype=2)
    assert B.lib().bcm_abstract_solver_c(None, C.byref(p), *cbs) == B.E_ARG
