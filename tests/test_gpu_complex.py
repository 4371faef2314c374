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


@pytest.mark.parametrize("nrhs", [1, 3])
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
