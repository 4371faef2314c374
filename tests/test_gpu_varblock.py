"""Block partitions the tuned kernels are not instantiated for: uniform block
sizes outside {3, 6, 8, 9, 12, 16, 24, 32} and variable partitions (one size
per block row, as CSPRNDS gives them, smatrix.h:1040-1048), through the C ABI
against the CPU oracle: products, sweeps, and full PCG / BiCGStab solves.

Parity bar: operators within 1e-13 relative; PCG the oracle's iteration count
exactly, BiCGStab within one; solutions within 1e-8 and the FP64 true
relative residual within the tolerance.
"""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def grid_blocks(B, nx, ny, nz, sizes, seed, symmetric, skew=0.3):
    """7-point block system on an nx x ny x nz grid with the given block size
    per grid point (diagonally dominant; symmetric -> upper storage, SPD)."""
    rng = np.random.default_rng(seed)
    ns = nx * ny * nz
    sp = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = lambda i, j, k: (i * ny + j) * nz + k   # noqa: E731
    nb = {}
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                r = idx(i, j, k)
                for di, dj, dk in ((1, 0, 0), (0, 1, 0), (0, 0, 1)):
                    if i + di < nx and j + dj < ny and k + dk < nz:
                        c = idx(i + di, j + dj, k + dk)
                        blk = rng.standard_normal((sizes[r], sizes[c])) / np.sqrt(sizes[r] + sizes[c])
                        nb[(r, c)] = blk
                        nb[(c, r)] = blk.T if symmetric else blk.T + skew * rng.standard_normal(blk.T.shape) / 3
    rowsum = np.zeros(sp[-1])
    for (r, c), blk in nb.items():
        rowsum[sp[r]:sp[r + 1]] += np.abs(blk).sum(axis=1)
    for r in range(ns):
        b = sizes[r]
        m = rng.standard_normal((b, b))
        d = m @ m.T / b + np.diag(rowsum[sp[r]:sp[r + 1]] + 1.0)
        nb[(r, r)] = d
    ia, ja, vals = [0], [], []
    for r in range(ns):
        cols = sorted(c for (rr, c) in nb if rr == r and (not symmetric or c >= r))
        for c in cols:
            ja.append(c)
            vals.append(nb[(r, c)].ravel(order="F"))
        ia.append(len(ja))
    m = B.Matrix.from_arrays(sp, ia, ja, np.concatenate(vals), symmetric=symmetric)
    rhs = rng.standard_normal(sp[-1])
    return m, rhs


def ordered(B, O, m, rhs, symmetric):
    order = m.rcm_order()
    mo = m.reorder(order)
    A = O.Block(*mo.arrays())
    if symmetric:
        f = B.Factor.ic2(mo)
        L = U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    else:
        f = B.Factor.ilu2(mo)
        L = O.Block(A.sprnd, *f.arrays(0, A.nsup))
        U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    return mo, f, A, L, U, mo.reorder_vector(order, rhs)


CASES = {
    "uniform-5": lambda n: [5] * n,
    "uniform-1": lambda n: [1] * n,
    "uniform-40": lambda n: [40] * n,
    "variable-2-7": lambda n: list(np.random.default_rng(1).choice([2, 3, 5, 7], n)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_products_and_sweeps(B, O, ctx, case):
    sizes = CASES[case](4 * 3 * 3)
    m, rhs = grid_blocks(B, 4, 3, 3, sizes, 5, symmetric=False)
    mo, f, A, L, U, b = ordered(B, O, m, rhs, False)
    print(f"varblock {case} n={A.n}", flush=True)
    x = np.random.default_rng(2).standard_normal((A.n, 3))
    dx = B.DeviceArray.from_host(np.asfortranarray(x))
    dy = B.DeviceArray((A.n, 3))
    ctx.mvm(mo, dx, dy, 3)
    ctx.sync()
    assert rel(dy.host(), O.mvma(A, x, 3)) < 1e-13
    ctx.mvm_t(mo, dx, dy, 3)
    ctx.sync()
    assert rel(dy.host(), O.mvmat(A, x, 3)) < 1e-13
    for which, fac in ((0, L), (1, U)):
        ctx.solve_l(f, which, dx, dy, 3)
        ctx.sync()
        assert rel(dy.host(), O.solvel(fac, x, 3)) < 1e-13
        ctx.solve_u(f, which, dx, dy, 3)
        ctx.sync()
        assert rel(dy.host(), O.solveu(fac, x, 3)) < 1e-13
    for prec, tol in ((0, 1e-13), (1, 2e-5), (2, 2e-2)):
        ref = O.solveu(U, O.solvel(L, x[:, 0]))
        ctx.precond(f, B.DeviceArray.from_host(x[:, 0].copy()), dy, precision=prec)
        ctx.sync()
        assert rel(dy.host()[:, 0] if dy.host().ndim == 2 else dy.host(), ref) < tol


@pytest.mark.parametrize("case", ["uniform-5", "variable-2-7"])
def test_solvers_on_variable_blocks(B, O, ctx, case):
    sizes = CASES[case](5 * 4 * 3)
    # nonsymmetric: ILU2 + BiCGStab
    m, rhs = grid_blocks(B, 5, 4, 3, sizes, 7, symmetric=False)
    mo, f, A, L, U, b = ordered(B, O, m, rhs, False)
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray(A.n)
    dx.zero()
    st = ctx.solve(mo, f, db, dx, B.default_opts(kind=B.BICGSTAB, eps=1e-10))
    xo, r = O.bicgstab(A, L, U, b, eps=1e-10)
    assert st.converged and st.relres_true <= 1e-10 and abs(st.iterations - r.nitperf) <= 1
    assert rel(dx.host(), xo) < 1e-8
    # symmetric (upper storage): IC2 + PCG, the oracle's iteration count exactly
    ms, rs = grid_blocks(B, 5, 4, 3, sizes, 8, symmetric=True)
    mo, f, A, L, U, b = ordered(B, O, ms, rs, True)
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray(A.n)
    dx.zero()
    st = ctx.solve(mo, f, db, dx, B.default_opts(kind=B.PCG, eps=1e-10, ittype=3, check_every=3))
    xo, r = O.cg(A, U, b, eps=1e-10, ittype=3)
    assert st.converged and st.iterations == r.updates
    assert rel(dx.host(), xo) < 1e-8 and st.relres_true <= 1e-10
