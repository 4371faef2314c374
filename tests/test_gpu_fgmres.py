"""Flexible GMRES(m) with an FP64 / FP32 / BF16 preconditioner against the oracle.

Parity bar: with the FP64 preconditioner the GPU solve matches the oracle's
FGMRES (modified Gram-Schmidt; the GPU uses CGS2, so the Arnoldi rounding
differs) to a relative solution difference <= 1e-8 and +-1 step; with the
FP32 and BF16 factors the solve still reaches FP64 true relative residual
<= 1e-10 and its solution differs from the FP64 oracle solution by <= 1e-8.
"""

import numpy as np
import pytest

from conftest import rel
from helpers import heat, seepage

pytestmark = pytest.mark.gpu


def run(B, ctx, mo, f, b, x0=None, **opts):
    db = B.DeviceArray.from_host(b)
    dx = B.DeviceArray.from_host(x0) if x0 is not None else B.DeviceArray(len(b))
    if x0 is None:
        dx.zero()
    st = ctx.solve(mo, f, db, dx, B.default_opts(kind=B.FGMRES, **opts))
    return dx.host(), st


@pytest.fixture(scope="module")
def sys3d(B, O):
    return seepage(B, O, 6, 5, 4, 24, seed=21, tau1=0.1, tau2=0.01)  # weak factor: ~11 steps


@pytest.mark.parametrize("check_every", [1, 4])
def test_fgmres_fp64_matches_oracle(B, O, ctx, sys3d, check_every):
    s = sys3d
    print(f"fgmres fp64 n={s['A'].n} check_every={check_every}", flush=True)
    x, st = run(B, ctx, s["mo"], s["f"], s["b"], eps=1e-10, restart=30, check_every=check_every)
    xo, r = O.fgmres(s["A"], s["L"], s["U"], s["b"], m=30, eps=1e-10)
    assert st.converged == 1 and st.breakdown == 0
    assert st.relres_true <= 1e-10
    assert abs(st.iterations - r.nitperf) <= 1, (st.iterations, r.nitperf)
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("prec", ["F32", "BF16"])
def test_fgmres_mixed_precision_reaches_fp64_tolerance(B, O, ctx, sys3d, prec):
    s = sys3d
    print(f"fgmres {prec} n={s['A'].n}", flush=True)
    x, st = run(B, ctx, s["mo"], s["f"], s["b"], eps=1e-10, restart=30, precision=getattr(B, prec))
    xo, r = O.fgmres(s["A"], s["L"], s["U"], s["b"], m=30, eps=1e-10)
    assert st.converged == 1 and st.relres_true <= 1e-10
    assert rel(x, xo) < 1e-8
    # the low-precision factors cost at most a few more outer steps
    assert st.iterations <= r.nitperf + (4 if prec == "F32" else 12)


def test_fgmres_short_restart_matches_oracle(B, O, ctx, sys3d):
    s = sys3d
    x, st = run(B, ctx, s["mo"], s["f"], s["b"], eps=1e-10, restart=4, check_every=3)
    xo, r = O.fgmres(s["A"], s["L"], s["U"], s["b"], m=4, eps=1e-10)
    assert st.converged == 1 and st.relres_true <= 1e-10
    assert st.restarts >= 2
    assert abs(st.iterations - r.nitperf) <= 2, (st.iterations, r.nitperf)
    assert rel(x, xo) < 1e-8


def test_fgmres_symmetric_ic2(B, O, ctx):
    h = heat(B, O, 14, 12, 9, seed=22)
    x, st = run(B, ctx, h["mo"], h["f"], h["b"], eps=1e-10, restart=20)
    xo, r = O.fgmres(h["A"], h["U"], h["U"], h["b"], m=20, eps=1e-10, symm=True)
    assert st.converged == 1 and st.relres_true <= 1e-10
    assert abs(st.iterations - r.nitperf) <= 1
    assert rel(x, xo) < 1e-8


def test_fgmres_warm_start_iteration_limit_and_zero_rhs(B, O, ctx, sys3d):
    s = sys3d
    xo, _ = O.fgmres(s["A"], s["L"], s["U"], s["b"], m=30, eps=1e-12)
    x, st = run(B, ctx, s["mo"], s["f"], s["b"], x0=xo * (1 + 1e-7), eps=1e-10)
    assert st.converged == 1 and st.iterations <= 8
    x, st = run(B, ctx, s["mo"], s["f"], s["b"], eps=1e-14, niter=2)
    assert st.converged == 0 and st.iterations == 2
    z = np.zeros_like(s["b"])
    x, st = run(B, ctx, s["mo"], s["f"], z, eps=1e-10)
    assert np.all(x == 0) and st.iterations == 0


def test_fgmres_restart_bounds_are_checked(B, ctx, sys3d):
    s = sys3d
    with pytest.raises(B.BcmError):
        run(B, ctx, s["mo"], s["f"], s["b"], restart=65)
