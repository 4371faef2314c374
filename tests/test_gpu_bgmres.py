"""GPU block GMRES (CSVector::BlockGMRES, gmres.cpp:337-1038) through the C ABI
against the oracle's restatement.

Parity bar: the same number of block steps as the oracle (+-1: the tall
reductions add in a different order), solutions within 1e-8 relative, the
recursion's per-step residual norms within 1e-6 relative (plus 1e-9 of the
initial residual), and with a fixed number of steps (eps = 0) the iterate of
the oracle after exactly those steps.
"""

import numpy as np
import pytest

from conftest import rel
from helpers import heat, seepage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys_ns(B, O):
    return seepage(B, O, 6, 5, 4, 24, seed=37)


def rhs_block(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal((n, nrhs)))


def run(B, ctx, s, b, x0=None, **opts):
    db = B.DeviceArray.from_host(b)
    dx = B.DeviceArray.from_host(x0 if x0 is not None else np.zeros_like(b))
    o = B.default_opts(kind=B.BLOCKGMRES, **opts)
    st, resi, hist = ctx.solve_block(s["mo"], s["f"], db, dx, b.shape[1], o)
    return dx.host(), st, resi, hist


def check_hist(hg, ho, steps, r0n):
    for k in range(steps):
        assert np.all(np.abs(hg[k] - ho[k]) <= 1e-6 * ho[k] + 1e-9 * r0n), (k, hg[k], ho[k])


@pytest.mark.parametrize("nrhs", [1, 3, 6, 8])
def test_block_gmres_matches_oracle(B, O, ctx, sys_ns, nrhs):
    s = sys_ns
    n = s["A"].n
    b = rhs_block(n, nrhs, 100 + nrhs)
    b[:, 0] = s["b"]
    print(f"block gmres nrhs={nrhs} n={n}", flush=True)
    x, st, resi, hist = run(B, ctx, s, b, niter=200, eps=1e-10, ittype=0, check_every=4)
    xo, r, ho, fo = O.block_gmres(s["A"], s["L"], s["U"], b, niter=200, eps=1e-10, sttype=1)
    assert r.updates == 1 and st.converged == 1 and st.breakdown == 0
    assert abs(st.iterations - r.nitperf) <= 1, (st.iterations, r.nitperf)
    assert rel(x, xo) < 1e-8
    r0n = np.linalg.norm(b, axis=0)
    check_hist(hist, ho, min(st.iterations, r.nitperf), r0n)
    Ad = s["mo"].dense()
    true = np.linalg.norm(b - Ad @ x, axis=0)
    assert np.allclose(resi, true, rtol=1e-6, atol=1e-12 * r0n.max())
    assert np.all(true <= 1.5e-10 * r0n)


def test_block_gmres_fixed_steps_and_absolute_test(B, O, ctx, sys_ns):
    """eps = 0: exactly niter steps and the oracle's iterate after them;
    ittype 1: the absolute test of sttype 0."""
    s = sys_ns
    n = s["A"].n
    b = rhs_block(n, 4, 7)
    x0 = 0.01 * rhs_block(n, 4, 8)
    for k in (1, 3, 7):
        x, st, resi, hist = run(B, ctx, s, b, x0, niter=k, eps=0.0, ittype=0, check_every=3)
        xo, r, ho, fo = O.block_gmres(s["A"], s["L"], s["U"], b, x0, niter=k, eps=0.0, sttype=1)
        assert st.iterations == k and st.converged == 0
        assert rel(x - x0, xo - x0) < 1e-9, k
        check_hist(hist, ho, k, np.linalg.norm(b, axis=0))
    x, st, resi, hist = run(B, ctx, s, b, x0, niter=100, eps=1e-6, ittype=1)
    xo, r, ho, fo = O.block_gmres(s["A"], s["L"], s["U"], b, x0, niter=100, eps=1e-6, sttype=0)
    assert st.converged == 1 and abs(st.iterations - r.nitperf) <= 1
    assert rel(x, xo) < 1e-6


def test_block_gmres_zero_column_and_relative_to_b(B, O, ctx, sys_ns):
    """A column with b = 0 and x0 = 0 is solved already (it stays zero and the
    block runs over the others); ittype 3 tests ||r_c|| < eps ||b_c||."""
    s = sys_ns
    n = s["A"].n
    b = rhs_block(n, 3, 11)
    b[:, 1] = 0.0
    x, st, resi, hist = run(B, ctx, s, b, niter=200, eps=1e-10, ittype=3)
    assert np.all(x[:, 1] == 0.0) and resi[1] == 0.0 and np.all(hist[:, 1] == 0.0)
    live = np.asfortranarray(b[:, [0, 2]])
    xo, r, ho, fo = O.block_gmres(s["A"], s["L"], s["U"], live, niter=200, eps=1e-10, sttype=1)
    assert st.converged == 1 and abs(st.iterations - r.nitperf) <= 1
    assert rel(x[:, [0, 2]], xo) < 1e-8
    assert np.all(resi[[0, 2]] <= 1.5e-10 * np.linalg.norm(live, axis=0))


def test_block_gmres_symmetric_ic2_and_single_column_kind(B, O, ctx):
    """Upper-stored symmetric matrix with an IC2 factor (SolveL and SolveU of
    the one factor), and bcm_solve with BCM_BLOCKGMRES = the one-column block."""
    h = heat(B, O, 14, 12, 9, seed=47)
    n = h["A"].n
    b = rhs_block(n, 2, 3)
    x, st, resi, hist = run(B, ctx, h, b, niter=150, eps=1e-10, ittype=0)
    xo, r, ho, fo = O.block_gmres(h["A"], h["U"], h["U"], b, niter=150, eps=1e-10, sttype=1, symm=True)
    assert st.converged == 1 and abs(st.iterations - r.nitperf) <= 1
    assert rel(x, xo) < 1e-8
    db = B.DeviceArray.from_host(b[:, 0].copy())
    dx = B.DeviceArray(n)
    dx.zero()
    st1 = ctx.solve(h["mo"], h["f"], db, dx, B.default_opts(kind=B.BLOCKGMRES, niter=150, eps=1e-10, ittype=0))
    x1o, r1, _, _ = O.block_gmres(h["A"], h["U"], h["U"], b[:, :1], niter=150, eps=1e-10, sttype=1, symm=True)
    assert st1.converged == 1 and abs(st1.iterations - r1.nitperf) <= 1
    assert rel(dx.host(), x1o[:, 0]) < 1e-8


def test_block_gmres_argument_errors(B, ctx, sys_ns):
    s = sys_ns
    n = s["A"].n
    b = rhs_block(n, 9, 1)
    db = B.DeviceArray.from_host(b)
    dx = B.DeviceArray(b.shape)
    with pytest.raises(B.BcmError):
        ctx.solve_block(s["mo"], s["f"], db, dx, 9, B.default_opts(kind=B.BLOCKGMRES))
