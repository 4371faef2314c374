"""Block Lanczos on the GPU (CSVector::BlockLanczosRight / Left / Central,
lanczos.cpp:932-3225) through the C ABI, against the oracle's restatement
(itself pinned to the reference's own compiled BlockLanczos* to 1e-12).

Parity bar: the same solution to 1e-8, an FP64 true relative residual within
the tolerance's reach, and the iteration at which the residual test first
passes within one test period (ichk): the GPU factors its blocks by
CholeskyQR2 where the reference uses Householder QR, which changes the
iterates only by rounding.
"""

import numpy as np
import pytest

from conftest import rel
from helpers import seepage

pytestmark = pytest.mark.gpu

KINDS = {0: "BLZ_RIGHT", 1: "BLZ_LEFT", 2: "BLZ_CENTRAL"}


@pytest.fixture(scope="module")
def sys24(B, O):
    return seepage(B, O, 6, 5, 4, 24, seed=61)


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["right", "left", "central"])
@pytest.mark.parametrize("nrhs", [1, 3, 6])
def test_block_lanczos_matches_oracle(B, O, ctx, sys24, variant, nrhs):
    s = sys24
    n = s["A"].n
    b = np.random.default_rng(70 + nrhs).standard_normal((n, nrhs))
    eps = 1e-10
    print(f"block lanczos variant={variant} nrhs={nrhs} n={n}", flush=True)
    db = B.DeviceArray.from_host(np.asfortranarray(b))
    dx = B.DeviceArray((n, nrhs))
    dx.zero()
    o = B.default_opts(kind=getattr(B, KINDS[variant]), eps=eps, niter=300, ichk=5)
    st, resi, _ = ctx.solve_block(s["mo"], s["f"], db, dx, nrhs, o)
    x = dx.host()
    xo, r = O.block_lanczos(variant, s["A"], s["L"], s["U"], b, eps=eps, niter=300, ichk=5)
    assert r.updates == 1 and st.converged == 1 and st.breakdown == 0
    assert abs(st.iterations - r.nitperf) <= 5, (st.iterations, r.nitperf)
    assert rel(x, xo) < 1e-8
    A = s["mo"].dense()
    for c in range(nrhs):
        rc = np.linalg.norm(b[:, c] - A @ x[:, c])
        assert abs(rc - resi[c]) <= 1e-8 * np.linalg.norm(b[:, c])
        assert rc <= 10 * eps * np.linalg.norm(b[:, c])


def test_block_lanczos_zero_column_and_warm_start(B, O, ctx, sys24):
    """A zero column (b = 0, x0 = 0: r0 = 0, which the reference's Solver
    leaves out of the block, solvers.cpp:1757-1771) stays zero while the
    other columns are solved; a warm start from a near solution converges."""
    s = sys24
    n = s["A"].n
    A = s["mo"].dense()
    b = np.random.default_rng(81).standard_normal((n, 3))
    b[:, 1] = 0.0
    db = B.DeviceArray.from_host(np.asfortranarray(b))
    dx = B.DeviceArray((n, 3))
    dx.zero()
    o = B.default_opts(kind=B.BLZ_CENTRAL, eps=1e-10, niter=300, ichk=5)
    st, resi, _ = ctx.solve_block(s["mo"], s["f"], db, dx, 3, o)
    x = dx.host()
    assert st.converged == 1
    assert np.all(x[:, 1] == 0.0)
    for c in (0, 2):
        assert np.linalg.norm(b[:, c] - A @ x[:, c]) <= 1e-9 * np.linalg.norm(b[:, c])
    # warm start: x0 within 1e-6 of the solution, so r0 = 1e-6 b; the test
    # is relative to r0 (lanczos.cpp:1536-1585), eps 1e-3 asks for 1e-9 |b|
    xs = np.linalg.solve(A, b[:, [0]])
    dx1 = B.DeviceArray.from_host(np.asfortranarray(xs * (1 + 1e-6)))
    db1 = B.DeviceArray.from_host(np.asfortranarray(b[:, [0]]))
    o1 = B.default_opts(kind=B.BLZ_CENTRAL, eps=1e-3, niter=300, ichk=5)
    st1, _, _ = ctx.solve_block(s["mo"], s["f"], db1, dx1, 1, o1)
    x1 = dx1.host()[:, 0]
    assert st1.converged == 1 and st1.iterations <= 30
    assert np.linalg.norm(b[:, 0] - A @ x1) <= 1e-9 * np.linalg.norm(b[:, 0])


@pytest.mark.parametrize("ittype", [2, 3, 4], ids=["central", "right", "left"])
def test_dropin_abstract_solver_block_lanczos(B, O, ittype):
    """AbstractSolver with the reference's default ittype 2 (and 3, 4): the
    block Lanczos variant Solver(RS) dispatches (solvers.cpp:1798-1800)."""
    import ctypes as C

    from test_gpu_solvers import _callbacks

    m, rhs = B.gen_seepage3d(4, 4, 3, 6, 0.3, 32)
    x = np.zeros(m.n)
    cbs = _callbacks(B, m, rhs, x)
    p = B.default_params(ittype=ittype, eps=1e-10, niter=500)
    code = B.lib().bcm_abstract_solver(None, C.byref(p), *cbs)
    assert code == 0, B.lib().bcm_last_error()
    xd = np.linalg.solve(m.dense(), rhs)
    assert rel(x, xd) < 1e-8
    assert p.nitperf > 0
