"""bcm_precond_noalloc (no counterpart in the reference: for callers that
cannot allocate device memory) against bcm_precond, through the C ABI.

Both apply the same factor sweeps (mvm.cpp:3140-3217, 3410-3484) under the
context's sweep schedule; the allocation-free variant keeps its intermediate
vector (and the sync-free control words) in a fixed per-GPU array. The bar
is bit-for-bit equality, on two calls in a row (the second on a smaller
system, so the array keeps the first call's contents), in each precision,
direction and schedule.
"""

import numpy as np
import pytest

from helpers import heat, seepage

pytestmark = pytest.mark.gpu


def _pair(B, ctx, f, r, precision, transpose):
    n, nrhs = r.shape
    dr = B.DeviceArray.from_host(np.asfortranarray(r))
    z1, z2 = B.DeviceArray((n, nrhs)), B.DeviceArray((n, nrhs))
    ctx.precond(f, dr, z1, nrhs=nrhs, precision=precision, transpose=transpose)   # also uploads the image
    ctx.precond_noalloc(f, dr, z2, nrhs=nrhs, precision=precision, transpose=transpose)
    ctx.sync()
    return z1.host(), z2.host()


@pytest.fixture(scope="module", params=[0, 1], ids=["level", "syncfree"])
def mctx(B, request):
    c = B.Context()
    c.set_sweep_mode(request.param)
    return c


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["f64", "f32", "bf16"])
@pytest.mark.parametrize("transpose", [False, True], ids=["M", "MT"])
def test_noalloc_matches_precond(B, O, mctx, precision, transpose):
    big = seepage(B, O, 6, 5, 4, 24, seed=71)        # N = 2880
    small = seepage(B, O, 3, 3, 2, 24, seed=72)      # N = 432
    rng = np.random.default_rng(73)
    print(f"noalloc precision={precision} transpose={transpose}", flush=True)
    for sysm, nrhs in ((big, 3), (small, 1)):
        r = rng.standard_normal((sysm["A"].n, nrhs))
        z1, z2 = _pair(B, mctx, sysm["f"], r, precision, transpose)
        assert np.isfinite(z1).all()
        assert np.array_equal(z1, z2), (precision, transpose, nrhs)


def test_noalloc_symmetric_factor(B, O, ctx):
    h = heat(B, O, 12, 10, 9, seed=74)
    r = np.random.default_rng(75).standard_normal((h["A"].n, 2))
    z1, z2 = _pair(B, ctx, h["f"], r, 0, False)
    assert np.array_equal(z1, z2)
    # and against the oracle's U^T then U sweeps (cg.cpp:146-147)
    ref = O.solveu(h["U"], O.solvel(h["U"], r[:, 0]))
    assert np.linalg.norm(z2[:, 0] - ref) <= 1e-13 * np.linalg.norm(ref)


def test_noalloc_rejects_what_it_cannot_do(B, O, ctx):
    s = seepage(B, O, 3, 3, 2, 24, seed=76)
    n = s["A"].n
    f = B.Factor.ilu2(s["mo"])          # fresh factor: no device image yet
    r = B.DeviceArray.from_host(np.ones(n))
    z = B.DeviceArray(n)
    with pytest.raises(B.BcmError) as e:
        ctx.precond_noalloc(f, r, z)
    assert e.value.code == B.E_ARG and "device image" in str(e.value)
    ctx.precond(f, r, z)
    ctx.sync()
    # n * nrhs beyond the 4,194,304-value scratch: refused before any launch
    nrhs = 4194304 // n + 1
    rb, zb = B.DeviceArray((n, nrhs)), B.DeviceArray((n, nrhs))
    with pytest.raises(B.BcmError) as e:
        ctx.precond_noalloc(f, rb, zb, nrhs=nrhs)
    assert e.value.code == B.E_ARG and "scratch" in str(e.value)
