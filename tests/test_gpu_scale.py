"""The solve phase at the configurations' own sizes, against the oracle.

Parity bar as in the smaller tests: PCG iteration counts equal to the
oracle's Cg (cg.cpp:31-349 restated in the oracle), solutions within 1e-8,
FP64 true relative residual within the requested tolerance; the SpMV within
1e-13 of MvmA (mvm.cpp:2530-2569). These sizes reach code paths the small
systems do not: SpMV waves that own several block rows (more than 8192 block
rows with one right-hand side), sync-free sweeps whose waves own several
rows in turn, the bounded-wait recovery of the single-system solve, uneven
cells in one batch, and the learned initial guess over 16^3 cells.
"""

from concurrent.futures import ThreadPoolExecutor

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


def elastic(B, O, m, prm, order=None, oracle=True):
    r, lc, nu, sh = prm
    mat, rhs = B.gen_elastic_cell(m, r, 10 ** lc, nu, sh, 20261205)
    order = mat.rcm_order() if order is None else order
    mo = mat.reorder(order)
    f = B.Factor.ic2(mo, B.default_params(tau1=0.1, tau2=0.01))
    b = mo.reorder_vector(order, rhs)
    out = dict(mo=mo, f=f, b=b, order=order)
    if oracle:
        out["A"] = O.Block(*mo.arrays())
        out["U"] = O.Block(out["A"].sprnd, *f.arrays(1, out["A"].nsup))
    return out


def oracle_cg_all(O, cells, eps=1e-10):
    """The oracle's Cg for every (cell, load case), on host threads."""
    jobs = [(c, q) for c in range(len(cells)) for q in range(6)]
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(lambda cq: O.cg(cells[cq[0]]["A"], cells[cq[0]]["U"], cells[cq[0]]["b"][:, cq[1]],
                                          eps=eps, ittype=3), jobs))
    out = [[None] * 6 for _ in cells]
    for (c, q), (x, r) in zip(jobs, res):
        out[c][q] = (x, r.updates)
    return out


def test_heat_64x64_matches_oracle(B, O, ctx):
    """The CPU-runnable 2D heat configuration at full size: 64 x 64 cells,
    block 9, RCM + IC2(1e-2, 1e-4) + PCG to rel-res 1e-10."""
    h = heat(B, O, 64, 64, 9, seed=20261203)
    print(f"heat 64x64 n={h['A'].n}", flush=True)
    x, st = solve(B, ctx, h["mo"], h["f"], h["b"], kind=B.PCG, eps=1e-10, ittype=3)
    xo, r = O.cg(h["A"], h["U"], h["b"], eps=1e-10, ittype=3)
    assert st.converged == 1 and st.iterations == r.updates
    assert rel(x, xo) < 1e-8 and st.relres_true <= 1e-10


def test_spmv_and_pcg_beyond_8192_block_rows(B, O, ctx):
    """9216 block rows: each SpMV wave owns two rows (alone and with the
    fused dot inside PCG)."""
    s = seepage(B, O, 24, 24, 16, 6, seed=31)
    A = s["A"]
    assert A.nsup > 8192
    print(f"spmv nsup={A.nsup} n={A.n}", flush=True)
    x = np.random.default_rng(3).standard_normal(A.n)
    dx, dy = dev(B, x), B.DeviceArray(A.n)
    ctx.mvm(s["mo"], dx, dy)
    ctx.sync()
    assert rel(dy.host(), O.mvma(A, x, 1)) < 1e-13
    h = heat(B, O, 96, 96, 9, seed=5)
    assert h["A"].nsup > 8192
    print(f"pcg heat 96x96 nsup={h['A'].nsup}", flush=True)
    x, st = solve(B, ctx, h["mo"], h["f"], h["b"], kind=B.PCG, eps=1e-10, ittype=3)
    xo, r = O.cg(h["A"], h["U"], h["b"], eps=1e-10, ittype=3)
    assert st.converged == 1 and st.iterations == r.updates
    assert rel(x, xo) < 1e-8


@pytest.fixture(scope="module")
def cell16(B, O):
    return elastic(B, O, 16, (0.3, 2.0, 0.3, 0.05))


@pytest.mark.parametrize("waves", [1, 8])
def test_pcg_16cube_cell_waves_own_several_rows(B, O, cell16, waves):
    """A 16^3 elasticity cell (4096 block rows of 24): with 1 or 8 resident
    waves per CU every wave of the sync-free sweeps owns several rows in
    turn (16 or 2); same iterations and solution as the oracle's Cg."""
    c = cell16
    ctx = B.Context()
    ctx.set_sweep_waves(waves)
    print(f"pcg 16^3 waves_per_cu={waves} n={c['A'].n}", flush=True)
    x, st = solve(B, ctx, c["mo"], c["f"], c["b"][:, 0], kind=B.PCG, eps=1e-10, ittype=3)
    xo, r = O.cg(c["A"], c["U"], c["b"][:, 0], eps=1e-10, ittype=3)
    assert st.converged == 1 and st.iterations == r.updates
    assert rel(x, xo) < 1e-8 and st.relres_true <= 1e-10


def test_wait_bound_recovery_returns_the_solution(B, O, cell16):
    """With the wait-bound flag raised from the host, every dependency wait
    of the first attempt exits at once; the solve detects it and repeats
    from x0 with cooperative launches (krylov.hip), returning the same
    solution and iteration count as an undisturbed solve."""
    c = cell16
    ctx = B.Context()
    x_ref, st_ref = solve(B, ctx, c["mo"], c["f"], c["b"][:, 1], kind=B.PCG, eps=1e-10, ittype=3, niter=300)
    ctx.raise_wait_flag()
    print("pcg 16^3 after a raised wait flag", flush=True)
    x, st = solve(B, ctx, c["mo"], c["f"], c["b"][:, 1], kind=B.PCG, eps=1e-10, ittype=3, niter=300)
    assert st.converged == 1 and st.iterations == st_ref.iterations
    assert np.array_equal(x, x_ref)
    # the flag is cleared: the next solve runs undisturbed
    x2, st2 = solve(B, ctx, c["mo"], c["f"], c["b"][:, 1], kind=B.PCG, eps=1e-10, ittype=3, niter=300)
    assert np.array_equal(x2, x_ref)


def test_batch_eight_uneven_cells(B, O, ctx):
    """Eight cells of 8^3..12^3 in one batch, FP64 and FP32 factors: every
    pair converges to rel-res 1e-10; three cells checked pair by pair
    against the oracle's Cg."""
    sizes = [8, 12, 9, 11, 10, 8, 12, 9]
    rng = np.random.default_rng(77)
    cells = [elastic(B, O, m, (rng.uniform(0.1, 0.45), rng.uniform(0, 3), rng.uniform(0.2, 0.45),
                                rng.uniform(-0.15, 0.15))) for m in sizes]
    bt = B.Batch([(c["mo"], c["f"]) for c in cells], prec_mask=0b011)
    b = bt.gather([c["b"] for c in cells])
    orc = oracle_cg_all(O, cells[:3])
    for prec in (B.F64, B.F32):
        print(f"batch of 8 uneven cells prec={prec} n={bt.n}", flush=True)
        db, dx = B.DeviceArray.from_host(b), B.DeviceArray(b.shape)
        dx.zero()
        st, its, rr = bt.solve(ctx, db, dx, B.default_opts(kind=B.BLOCKPCG, eps=1e-10, precision=prec), 6)
        assert st.converged == st.pairs == 48 and rr.max() <= 1e-10
        xs = bt.scatter(dx.host())
        for ci in range(3):
            for q in range(6):
                xo, ito = orc[ci][q]
                assert rel(xs[ci][:, q], xo) < 1e-8, (prec, ci, q)
                if prec == B.F64:
                    assert abs(int(its[ci, q]) - ito) <= 1, (ci, q, its[ci, q], ito)


def test_learned_guess_16cube_cells(B, O, ctx):
    """The learned initial guess at its configuration's cell size: 16^3
    elasticity cells sharing one template ordering, 48 training cells solved
    from zero, POD (16 modes per load case: fewer than the 48 snapshots of
    each) + BF16-MFMA MLP, then 4 unseen cells solved from the learned x0.
    Every pair reaches FP64 true rel-res 1e-10 and matches the oracle's Cg
    from zero within 1e-8; the guess is closer than zero and lowers the mean
    iteration count."""
    ntrain, ntest = 48, 4
    rng = np.random.default_rng(20261206)

    def draw(k):
        return np.stack([rng.uniform(0.1, 0.45, k), rng.uniform(0, 3, k), rng.uniform(0.2, 0.45, k),
                         rng.uniform(-0.15, 0.15, k)], axis=1)

    ptrain, ptest = draw(ntrain), draw(ntest)
    mat0, _ = B.gen_elastic_cell(16, 0.3, 10.0, 0.3, 0.0, 20261205)
    order = mat0.rcm_order()
    with ThreadPoolExecutor(8) as ex:
        train = list(ex.map(lambda p: elastic(B, O, 16, p, order, oracle=False), ptrain))
        test = list(ex.map(lambda p: elastic(B, O, 16, p, order), ptest))
    n = train[0]["mo"].n
    opts = B.default_opts(kind=B.BLOCKPCG, eps=1e-10, precision=B.F32)

    def bsolve(cs, x0=None):
        bt = B.Batch([(c["mo"], c["f"]) for c in cs], prec_mask=1 << B.F32)
        db = B.DeviceArray.from_host(bt.gather([c["b"] for c in cs]))
        dx = B.DeviceArray.from_host(x0) if x0 is not None else B.DeviceArray(db.shape)
        if x0 is None:
            dx.zero()
        st, its, rr = bt.solve(ctx, db, dx, opts, 6, x0 is not None)
        assert st.converged == st.pairs and rr.max() <= 1e-10
        return bt, dx, its

    g = B.Guess(n, 6, 4, nmodes=16, nhidden=64, capacity=ntrain, seed=11)
    for k0 in range(0, ntrain, 16):
        print(f"learned 16^3: training cells {k0}..{k0 + 15}", flush=True)
        _, dx, _ = bsolve(train[k0:k0 + 16])
        g.add(ctx, ptrain[k0:k0 + 16], dx)
    rep = g.fit(ctx, steps=3000, lr=3e-3)
    print(f"guess fit: energy {rep.pod_energy:.6f} loss {rep.loss_first:.3e} -> {rep.loss_final:.3e}", flush=True)
    assert rep.loss_final < 0.1 * rep.loss_first
    x0 = B.DeviceArray((n * ntest, 6))
    g.predict(ctx, ptest, x0)
    guess = x0.host()
    _, dxz, its0 = bsolve(test)
    exact = dxz.host()
    err = np.linalg.norm(guess - exact) / np.linalg.norm(exact)
    bt, dx, its1 = bsolve(test, guess)
    print(f"guess rel err {err:.3e}; iterations zero {its0.mean():.1f} learned {its1.mean():.1f}", flush=True)
    assert err < 0.5
    assert its1.mean() < its0.mean()
    xs = bt.scatter(dx.host())
    orc = oracle_cg_all(O, test)
    for ci in range(ntest):
        for q in range(6):
            assert rel(xs[ci][:, q], orc[ci][q][0]) < 1e-8, (ci, q)
