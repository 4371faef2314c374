"""Sync-free triangular sweeps (one launch per sweep; static row ownership;
the output pre-filled with a sentinel that marks unfinished rows; each row's
dependency values staged through LDS behind a bounded, wave-voted wait)
against the CPU oracle, and against the level-scheduled sweeps."""

import numpy as np
import pytest

from conftest import rel
from helpers import heat, seepage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fctx(B):
    c = B.Context()
    c.set_sweep_mode(1)
    return c


def dev(B, x):
    return B.DeviceArray.from_host(np.asfortranarray(x))


@pytest.mark.parametrize("nrhs", [1, 2, 5, 6])
def test_syncfree_ic2_sweeps(B, O, fctx, nrhs):
    h = heat(B, O, 9, 8, 9, seed=5)
    U = h["U"]
    print(f"syncfree ic2 nrhs={nrhs} nsup={U.nsup}", flush=True)
    x = np.random.default_rng(3).standard_normal((U.n, nrhs))
    dx, dy, dz = dev(B, x), B.DeviceArray((U.n, nrhs)), B.DeviceArray((U.n, nrhs))
    fctx.solve_l(h["f"], 1, dx, dy, nrhs)
    fctx.solve_u(h["f"], 1, dy, dz, nrhs)
    fctx.sync()
    y = O.solvel(U, x, nrhs)
    assert rel(dy.host(), y) < 1e-13
    assert rel(dz.host(), O.solveu(U, y, nrhs)) < 1e-13


def test_syncfree_ilu2_all_directions(B, O, fctx):
    s = seepage(B, O, 6, 5, 4, 24, seed=6)
    L, U = s["L"], s["U"]
    print(f"syncfree ilu2 nsup={L.nsup}", flush=True)
    x = np.random.default_rng(4).standard_normal(L.n)
    dx, dy = dev(B, x), B.DeviceArray(L.n)
    for which, fn, ref in ((0, fctx.solve_l, O.solvel(L, x)), (1, fctx.solve_l, O.solvel(U, x)),
                           (1, fctx.solve_u, O.solveu(U, x)), (0, fctx.solve_u, O.solveu(L, x))):
        fn(s["f"], which, dx, dy)
        fctx.sync()
        assert rel(dy.host(), ref) < 1e-13


@pytest.mark.parametrize("precision,tol", [(1, 2e-5), (2, 2e-2)])
def test_syncfree_low_precision(B, O, fctx, precision, tol):
    s = seepage(B, O, 5, 4, 3, 24, seed=7)
    r = np.random.default_rng(5).standard_normal(s["A"].n)
    ref = O.solveu(s["U"], O.solvel(s["L"], r))
    dr, dz = dev(B, r), B.DeviceArray(s["A"].n)
    fctx.precond(s["f"], dr, dz, precision=precision)
    fctx.sync()
    assert rel(dz.host(), ref) < tol


def test_syncfree_solvers(B, O, fctx):
    h = heat(B, O, 16, 14, 9, seed=11)
    db, dx = dev(B, h["b"]), B.DeviceArray(h["A"].n)
    dx.zero()
    st = fctx.solve(h["mo"], h["f"], db, dx, B.default_opts(kind=B.PCG, eps=1e-10, ittype=0))
    xo, r = O.cg(h["A"], h["U"], h["b"], eps=1e-10, ittype=0)
    assert st.converged and st.iterations == r.updates and rel(dx.host(), xo) < 1e-8
    s = seepage(B, O, 6, 5, 4, 24, seed=12)
    db, dx = dev(B, s["b"]), B.DeviceArray(s["A"].n)
    dx.zero()
    st = fctx.solve(s["mo"], s["f"], db, dx, B.default_opts(kind=B.BICGSTAB, eps=1e-10))
    xo, r = O.bicgstab(s["A"], s["L"], s["U"], s["b"], eps=1e-10)
    assert st.converged and st.relres_true <= 1e-10 and rel(dx.host(), xo) < 1e-8
