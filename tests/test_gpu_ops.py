"""GPU operators through the C ABI against the CPU oracle (same seeded inputs).

Tolerances: fp64 kernels differ from the sequential oracle only by the
summation order of each row (and FMA contraction): relative 1e-13. fp32
factors: 2e-5; bf16 factor storage with fp32 accumulation: 2e-2.
"""

import numpy as np
import pytest

from conftest import rel
from helpers import heat, seepage

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[0, 1, 2], ids=["levels", "syncfree", "syncfree-natural"])
def ctx(B, request):
    c = B.Context()
    c.set_sweep_mode(request.param)
    return c


def dev(B, x):
    return B.DeviceArray.from_host(np.asfortranarray(x))


@pytest.mark.parametrize("bla", [3, 6, 8, 12, 16, 24, 32])
@pytest.mark.parametrize("nrhs", [1, 2, 3, 5, 7])
def test_spmv_and_transpose(B, O, ctx, bla, nrhs):
    s = seepage(B, O, 4, 3, 3, bla, seed=bla)
    A = s["A"]
    print(f"spmv bla={bla} nrhs={nrhs} n={A.n}", flush=True)
    x = np.random.default_rng(bla).standard_normal((A.n, nrhs))
    dx, dy = dev(B, x), B.DeviceArray((A.n, nrhs))
    ctx.mvm(s["mo"], dx, dy, nrhs)
    ctx.sync()
    assert rel(dy.host(), O.mvma(A, x, nrhs)) < 1e-13
    ctx.mvm_t(s["mo"], dx, dy, nrhs)
    ctx.sync()
    assert rel(dy.host(), O.mvmat(A, x, nrhs)) < 1e-13


def test_spmv_symmetric_and_residual(B, O, ctx):
    h = heat(B, O, 9, 8, 9, seed=4)
    A = h["A"]
    x = np.random.default_rng(2).standard_normal(A.n)
    dx, dy = dev(B, x), B.DeviceArray(A.n)
    ctx.mvm(h["mo"], dx, dy)
    ctx.sync()
    ax = O.mvmasymm(A, x)
    assert rel(dy.host(), ax) < 1e-13
    db, dr = dev(B, h["b"]), B.DeviceArray(A.n)
    ctx.residual(h["mo"], db, dx, dr)
    ctx.sync()
    assert rel(dr.host(), h["b"] - ax) < 1e-13


def ragged_matrix(B, O, nsup, bla, maxblk, seed):
    """Random nonsymmetric block matrix whose rows hold 1..maxblk blocks (the
    diagonal always), so rows span several staging chunks of the products
    and sweeps, and some hold the diagonal block alone."""
    rng = np.random.default_rng(seed)
    ia, ja = [0], []
    for i in range(nsup):
        k = 1 if i % 7 == 3 else int(rng.integers(1, maxblk + 1))
        cols = set(rng.choice(nsup, size=min(k, nsup), replace=False).tolist()) | {i}
        ja.extend(sorted(cols))
        ia.append(len(ja))
    a = rng.standard_normal(len(ja) * bla * bla) / np.sqrt(bla * maxblk)
    off = 0
    for i in range(nsup):  # dominant diagonal blocks keep the factorization stable
        for j in range(ia[i], ia[i + 1]):
            if ja[j] == i:
                a[off:off + bla * bla] += 4.0 * np.eye(bla).reshape(-1)
            off += bla * bla
    m = B.Matrix.from_arrays(np.arange(nsup + 1) * bla, ia, ja, a, symmetric=False)
    return m, O.Block(*m.arrays())


@pytest.mark.parametrize("bla", [12, 16, 24])
@pytest.mark.parametrize("nrhs", [2, 3, 6])
def test_spmv_long_ragged_rows(B, O, ctx, bla, nrhs):
    """Rows of up to 22 blocks: several LDS staging chunks per row, rows
    shorter than the matrix prefetch, a diagonal-only row."""
    m, A = ragged_matrix(B, O, 48, bla, 22, seed=bla + nrhs)
    print(f"ragged spmv bla={bla} nrhs={nrhs} n={A.n}", flush=True)
    x = np.random.default_rng(bla).standard_normal((A.n, nrhs))
    dx, dy = dev(B, x), B.DeviceArray((A.n, nrhs))
    ctx.mvm(m, dx, dy, nrhs)
    ctx.sync()
    assert rel(dy.host(), O.mvma(A, x, nrhs)) < 1e-13
    ctx.mvm_t(m, dx, dy, nrhs)
    ctx.sync()
    assert rel(dy.host(), O.mvmat(A, x, nrhs)) < 1e-13


@pytest.mark.parametrize("nrhs", [1, 2, 6])
def test_sweeps_long_ragged_rows(B, O, ctx, nrhs):
    """ILU2 factors of a matrix with rows of up to 16 blocks: factor rows that
    span several dependency-staging rounds of the level sweep."""
    m, A = ragged_matrix(B, O, 40, 24, 16, seed=31 + nrhs)
    f = B.Factor.ilu2(m)
    L = O.Block(A.sprnd, *f.arrays(0, A.nsup))
    U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    print(f"ragged sweeps nrhs={nrhs} n={A.n}", flush=True)
    x = np.random.default_rng(12).standard_normal((A.n, nrhs))
    dx, dy, dz = dev(B, x), B.DeviceArray((A.n, nrhs)), B.DeviceArray((A.n, nrhs))
    ctx.solve_l(f, 0, dx, dy, nrhs)
    ctx.solve_u(f, 1, dy, dz, nrhs)
    ctx.sync()
    y = O.solvel(L, x, nrhs)
    assert rel(dy.host(), y) < 1e-12
    assert rel(dz.host(), O.solveu(U, y, nrhs)) < 1e-12


@pytest.mark.parametrize("nrhs", [1, 2, 6])
def test_sweeps_ic2(B, O, ctx, nrhs):
    h = heat(B, O, 9, 8, 9, seed=5)
    U = h["U"]
    x = np.random.default_rng(3).standard_normal((U.n, nrhs))
    dx, dy, dz = dev(B, x), B.DeviceArray((U.n, nrhs)), B.DeviceArray((U.n, nrhs))
    ctx.solve_l(h["f"], 1, dx, dy, nrhs)
    ctx.solve_u(h["f"], 1, dy, dz, nrhs)
    ctx.sync()
    y = O.solvel(U, x, nrhs)
    assert rel(dy.host(), y) < 1e-13
    assert rel(dz.host(), O.solveu(U, y, nrhs)) < 1e-13


def test_sweeps_ilu2_all_directions(B, O, ctx):
    s = seepage(B, O, 5, 4, 3, 24, seed=6)
    L, U = s["L"], s["U"]
    x = np.random.default_rng(4).standard_normal(L.n)
    dx, dy = dev(B, x), B.DeviceArray(L.n)
    for which, fn, ref in ((0, ctx.solve_l, O.solvel(L, x)), (1, ctx.solve_l, O.solvel(U, x)),
                           (1, ctx.solve_u, O.solveu(U, x)), (0, ctx.solve_u, O.solveu(L, x))):
        fn(s["f"], which, dx, dy)
        ctx.sync()
        assert rel(dy.host(), ref) < 1e-13


@pytest.mark.parametrize("precision,tol", [(0, 1e-13), (1, 2e-5), (2, 2e-2)])
def test_preconditioner_precisions(B, O, ctx, precision, tol):
    s = seepage(B, O, 5, 4, 3, 24, seed=7)
    r = np.random.default_rng(5).standard_normal(s["A"].n)
    ref = O.solveu(s["U"], O.solvel(s["L"], r))
    dr, dz = dev(B, r), B.DeviceArray(s["A"].n)
    ctx.precond(s["f"], dr, dz, precision=precision)
    ctx.sync()
    assert rel(dz.host(), ref) < tol
    # transposed preconditioner: L^-T U^-T = U.SolveL then L.SolveU (lanczos.cpp:264-265)
    reft = O.solveu(s["L"], O.solvel(s["U"], r))
    ctx.precond(s["f"], dr, dz, precision=precision, transpose=True)
    ctx.sync()
    assert rel(dz.host(), reft) < tol


def test_dot(B, ctx):
    x = np.random.default_rng(6).standard_normal(100003)
    y = np.random.default_rng(7).standard_normal(100003)
    assert abs(ctx.dot(100003, dev(B, x), dev(B, y)) - x @ y) < 1e-10 * np.abs(x * y).sum()
    assert ctx.dot(0, dev(B, x), dev(B, y)) == 0.0


def test_single_block_row(B, O, ctx):
    a = np.random.default_rng(8).standard_normal(64) + 8 * np.eye(8).reshape(-1)
    m = B.Matrix.from_arrays([0, 8], [0, 1], [0], a, symmetric=False)
    x = np.random.default_rng(9).standard_normal(8)
    dx, dy = dev(B, x), B.DeviceArray(8)
    ctx.mvm(m, dx, dy)
    ctx.sync()
    assert rel(dy.host(), a.reshape(8, 8).T @ x) < 1e-14
    f = B.Factor.ilu2(m)
    ctx.precond(f, dx, dy)
    ctx.sync()
    assert rel(dy.host(), np.linalg.solve(a.reshape(8, 8).T, x)) < 1e-12


def test_block_rows_beyond_64_are_an_error(B, ctx):
    """Blocks of up to 64 rows run on the device; larger ones are refused."""
    b = 65
    a = np.random.default_rng(10).standard_normal(b * b) + 20 * np.eye(b).reshape(-1)
    m = B.Matrix.from_arrays([0, b], [0, 1], [0], a, symmetric=False)
    dx, dy = B.DeviceArray(b), B.DeviceArray(b)
    with pytest.raises(B.BcmError) as e:
        ctx.mvm(m, dx, dy)
    assert e.value.code == B.E_ARG
