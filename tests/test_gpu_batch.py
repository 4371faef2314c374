"""Batched block PCG over uneven elasticity cells x 6 load cases against the oracle.

Parity bar: every (cell, load case) pair of the batched FP64 solve matches
the oracle's Cg on that cell alone (cg.cpp:31-349 restated in the oracle):
iteration count within one (the batched dots add chunk partials in a
different order) and relative solution difference <= 1e-8; FP32/BF16 factor
sweeps (flexible CG, FP64 outer recurrence) reach FP64 true relative
residual <= 1e-10 and stay within 1e-8 of the FP64 oracle solution. Eager
and hipGraph-replayed launches of a schedule give bit-identical results.
"""

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

SIZES = [3, 5, 4, 7, 6]  # uneven periodic cells (m^3 block rows of 24)


def cell(B, O, m, seed):
    rng = np.random.default_rng(seed)
    radius, contrast, nu = rng.uniform(0.1, 0.45), 10 ** rng.uniform(0, 3), rng.uniform(0.2, 0.45)
    mat, rhs = B.gen_elastic_cell(m, radius, contrast, nu, 0.0, 20261205 + m)
    order = mat.rcm_order()
    mo = mat.reorder(order)
    f = B.Factor.ic2(mo, B.default_params(tau1=0.1, tau2=0.01))  # the cell family's IC2 thresholds
    b = mo.reorder_vector(order, rhs)
    A = O.Block(*mo.arrays())
    U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    return dict(mo=mo, f=f, b=b, A=A, U=U)


@pytest.fixture(scope="module")
def cells(B, O):
    return [cell(B, O, m, 100 + i) for i, m in enumerate(SIZES)]


@pytest.fixture(scope="module")
def batch(B, cells):
    return B.Batch([(c["mo"], c["f"]) for c in cells], prec_mask=0b111)


@pytest.fixture(scope="module")
def oracle_sol(O, cells):
    out = []
    for c in cells:
        xs, its = [], []
        for q in range(6):
            x, r = O.cg(c["A"], c["U"], c["b"][:, q], eps=1e-10, ittype=3)
            xs.append(x)
            its.append(r.updates)
        out.append((np.stack(xs, axis=1), np.array(its)))
    return out


def run(B, ctx, bt, cells, nrhs=6, x0=None, fallback=False, galerkin=False, **opts):
    b = bt.gather([c["b"][:, :nrhs] for c in cells])
    db = B.DeviceArray.from_host(b)
    dx = B.DeviceArray.from_host(x0) if x0 is not None else B.DeviceArray(b.shape)
    if x0 is None:
        dx.zero()
    o = B.default_opts(kind=B.BLOCKPCG, **opts)
    st, its, rr = bt.solve(ctx, db, dx, o, nrhs, fallback, galerkin)
    return bt.scatter(dx.host()), st, its, rr


def test_batch_layout(B, batch, cells):
    assert batch.ncell == len(SIZES)
    assert batch.n == sum(24 * m ** 3 for m in SIZES)
    assert list(np.diff(batch.elem0)) == [24 * m ** 3 for m in SIZES]


def test_batch_fp64_matches_oracle_per_pair(B, ctx, batch, cells, oracle_sol):
    print(f"batch fp64 cells={SIZES} n={batch.n}", flush=True)
    xs, st, its, rr = run(B, ctx, batch, cells, eps=1e-10, precision=B.F64)
    assert st.converged == st.pairs == 6 * len(SIZES) and st.breakdowns == 0
    assert rr.max() <= 1e-10
    for c, (x, (xo, ito)) in enumerate(zip(xs, oracle_sol)):
        assert np.all(np.abs(its[c] - ito) <= 1), (c, its[c], ito)
        for q in range(6):
            assert rel(x[:, q], xo[:, q]) < 1e-8, (c, q)


@pytest.mark.parametrize("prec", ["F32", "BF16"])
def test_batch_mixed_precision(B, ctx, batch, cells, oracle_sol, prec):
    print(f"batch {prec}", flush=True)
    xs, st, its, rr = run(B, ctx, batch, cells, eps=1e-10, precision=getattr(B, prec))
    assert st.converged == st.pairs and rr.max() <= 1e-10
    for c, (x, (xo, ito)) in enumerate(zip(xs, oracle_sol)):
        assert rel(x, xo) < 1e-8, c
        assert np.all(its[c] <= ito + (3 if prec == "F32" else ito + 10)), (c, its[c], ito)


def test_batch_schedules_agree(B, ctx, batch, cells):
    """Graph replay and eager launches of one sweep schedule are bit-identical;
    the level-scheduled and sync-free sweeps add the same terms in different
    orders, so across them the solutions agree to rounding."""
    ref, first = {}, {}
    for mode, graph, every in [(1, 0, 8), (1, 1, 4), (1, 0, 1), (0, 0, 8), (0, 1, 16), (0, 0, 3)]:
        print(f"schedule mode={mode} graph={graph} every={every}", flush=True)
        batch.set_tune(sweep_mode=mode, graph=graph, check_every=every, waves_per_cu=8)
        for prec in (B.F64, B.F32):
            xs, st, its, rr = run(B, ctx, batch, cells, eps=1e-10, precision=prec)
            assert st.converged == st.pairs and rr.max() <= 1e-10
            x = np.concatenate([v.ravel() for v in xs])
            key = (mode, prec)
            if key in ref:
                assert np.array_equal(x, ref[key][0]) and np.array_equal(its, ref[key][1]), (mode, graph, every)
            else:
                ref[key] = (x, its.copy())
            if prec in first:
                assert rel(x, first[prec][0]) < 1e-8 and np.all(np.abs(its - first[prec][1]) <= 1)
            else:
                first[prec] = (x, its.copy())
    batch.set_tune(sweep_mode=1, graph=0, check_every=8, waves_per_cu=8)


def test_syncfree_wait_bound_recovery(B, ctx, batch, cells):
    """With the wait-bound flag raised from the host, every sync-free wait of
    the batched solve exits at once; the solve notices, repeats itself from
    x0 with level-scheduled sweeps and reports it: the result equals a
    level-scheduled solve bit for bit. The next solve runs undisturbed."""
    batch.set_tune(sweep_mode=0, graph=0, check_every=4, waves_per_cu=8)
    x_ref, st_ref, its_ref, _ = run(B, ctx, batch, cells, eps=1e-10, precision=B.F32, niter=300)
    assert st_ref.recovered == 0
    batch.set_tune(sweep_mode=1, graph=0, check_every=4, waves_per_cu=8)
    ctx.raise_wait_flag()
    print("batched sync-free solve after a raised wait flag", flush=True)
    x, st, its, rr = run(B, ctx, batch, cells, eps=1e-10, precision=B.F32, niter=300)
    assert st.recovered == 1 and st.converged == st.pairs and rr.max() <= 1e-10
    assert all(np.array_equal(a, b) for a, b in zip(x, x_ref)) and np.array_equal(its, its_ref)
    x2, st2, _, rr2 = run(B, ctx, batch, cells, eps=1e-10, precision=B.F32, niter=300)
    assert st2.recovered == 0 and st2.converged == st2.pairs and rr2.max() <= 1e-10
    batch.set_tune(sweep_mode=1, graph=0, check_every=8, waves_per_cu=8)


def test_graph_capture_while_another_thread_builds(B, ctx, cells):
    """Graph-replayed solves on this thread while another thread builds and
    drops batches (device allocation and frees, pinned staging copies, the
    symmetric expansion's helper stream): every capture excludes the other
    thread's set-up calls, and each solve (a fresh capture, since setting the
    schedule drops the graph) equals a quiet run bit for bit."""
    import threading

    mats = [(c["mo"], c["f"]) for c in cells]
    bt = B.Batch(mats, prec_mask=1 << B.F32)
    tune = dict(sweep_mode=0, graph=1, check_every=4, waves_per_cu=8)
    bt.set_tune(**tune)
    x_ref, st, its_ref, _ = run(B, ctx, bt, cells, eps=1e-10, precision=B.F32)
    assert st.converged == st.pairs
    x_ref = np.concatenate([v.ravel() for v in x_ref])
    stop, built, err = threading.Event(), [0], []

    def builder():
        try:
            while not stop.is_set():
                B.Batch(mats, prec_mask=1 << B.F32)  # built on this thread and dropped here
                built[0] += 1
        except BaseException as e:  # reported by the main thread
            err.append(e)

    th = threading.Thread(target=builder)
    th.start()
    try:
        for k in range(4):
            print(f"concurrent build + capture, solve {k}", flush=True)
            bt.set_tune(**tune)
            xs, st, its, rr = run(B, ctx, bt, cells, eps=1e-10, precision=B.F32)
            x = np.concatenate([v.ravel() for v in xs])
            assert st.converged == st.pairs
            assert np.array_equal(x, x_ref) and np.array_equal(its, its_ref), k
    finally:
        stop.set()
        th.join()
    assert not err, err
    assert built[0] >= 1


def test_batch_warm_start_fallback_and_zero_rhs(B, ctx, batch, cells, oracle_sol):
    good = batch.gather([xo * (1 + 1e-7) for xo, _ in oracle_sol])
    xs, st, its, rr = run(B, ctx, batch, cells, x0=good, fallback=True, eps=1e-10)
    assert st.converged == st.pairs
    for c, (xo, ito) in enumerate(oracle_sol):
        assert its[c].max() < ito.min(), (c, its[c], ito)
    bad = good * -50.0  # ||b - A x0|| >= ||b||: every pair falls back to x0 = 0
    xs, st, its2, rr = run(B, ctx, batch, cells, x0=bad, fallback=True, eps=1e-10)
    assert st.converged == st.pairs
    for c, (xo, ito) in enumerate(oracle_sol):
        assert np.all(np.abs(its2[c] - ito) <= 1)
    zero = [dict(c, b=np.zeros_like(c["b"])) for c in cells]
    xs, st, its3, rr = run(B, ctx, batch, zero, eps=1e-10)
    assert np.all(its3 == 0) and all(np.all(x == 0) for x in xs)


@pytest.mark.parametrize("nrhs", [1, 3])
def test_batch_fewer_rhs(B, ctx, batch, cells, oracle_sol, nrhs):
    xs, st, its, rr = run(B, ctx, batch, cells, nrhs=nrhs, eps=1e-10)
    assert st.pairs == nrhs * len(SIZES) and st.converged == st.pairs
    for c, (x, (xo, ito)) in enumerate(zip(xs, oracle_sol)):
        assert rel(x, xo[:, :nrhs]) < 1e-8


def test_batch_autotune_and_errors(B, ctx, batch, cells):
    b = batch.gather([c["b"] for c in cells])
    db = B.DeviceArray.from_host(b)
    t = batch.autotune(ctx, db, B.default_opts(kind=B.BLOCKPCG, precision=B.F32), 6)
    print(f"tuned: mode={t.sweep_mode} waves={t.waves_per_cu} graph={t.graph} every={t.check_every} "
          f"{t.ms_per_iter:.3f} ms/iter", flush=True)
    assert t.sweep_mode in (0, 1) and t.check_every in (4, 16) and t.ms_per_iter > 0
    dx = B.DeviceArray(b.shape)
    dx.zero()
    st, its, rr = batch.solve(ctx, db, dx, B.default_opts(kind=B.BLOCKPCG, precision=B.F32), 6)
    assert st.converged == st.pairs
    with pytest.raises(B.BcmError):
        batch.solve(ctx, db, dx, B.default_opts(kind=B.BLOCKPCG), 2)


def test_batch_galerkin_start(B, ctx, batch, cells, oracle_sol):
    """Start flag 2: each cell's six guesses X become X C with
    (X^T A X) C = X^T B. With niter = 0 the solve returns the start: it equals
    the dense computation per cell to 1e-9, and every column is no farther
    from the cell's solution in the A-norm than its own guess or than zero.
    A cell whose guesses are linearly dependent keeps them. Then full solves
    from the combined start converge to the oracle's solutions."""
    rng = np.random.default_rng(41)
    guesses = []
    for c, (xo, _) in enumerate(oracle_sol):
        g = xo * (1.0 + 0.2 * rng.standard_normal(xo.shape[1]))[None, :] + 0.02 * rng.standard_normal(xo.shape)
        if c == 1:
            g = -50.0 * g  # far worse than zero
        if c == 3:
            g[:, 4] = g[:, 2]  # dependent guesses: the cell keeps them
        guesses.append(g)
    x0 = batch.gather(guesses)
    print("batch galerkin start, niter 0", flush=True)
    xs, st, its, _ = run(B, ctx, batch, cells, x0=x0, galerkin=True, eps=1e-10, niter=0)
    for c, (cl, g, x) in enumerate(zip(cells, guesses, xs)):
        A = cl["mo"].dense()
        b = cl["b"]
        xo = oracle_sol[c][0]
        if c == 3:
            assert np.array_equal(x, g), "a cell with dependent guesses keeps them"
            continue
        G = g.T @ A @ g
        want = g @ np.linalg.solve(0.5 * (G + G.T), g.T @ b)
        assert rel(x, want) < 1e-9, c
        for q in range(6):
            def anorm(v):
                return float(np.sqrt(v @ A @ v))
            e_new, e_own, e_zero = anorm(x[:, q] - xo[:, q]), anorm(g[:, q] - xo[:, q]), anorm(xo[:, q])
            assert e_new <= min(e_own, e_zero) * (1 + 1e-10), (c, q, e_new, e_own, e_zero)
    print("batch galerkin start, full solves", flush=True)
    xs, st, its, rr = run(B, ctx, batch, cells, x0=x0, fallback=True, galerkin=True, eps=1e-10)
    assert st.converged == st.pairs and rr.max() <= 1e-10
    for x, (xo, _) in zip(xs, oracle_sol):
        assert rel(x, xo) < 1e-8
    import ctypes

    b = batch.gather([c["b"] for c in cells])
    db, dx = B.DeviceArray.from_host(b), B.DeviceArray.from_host(b)
    bst, o = B.BatchStats(), B.default_opts(kind=B.BLOCKPCG)
    code = B.lib().bcm_batch_solve(ctx.h, batch.h, ctypes.byref(o), 6, B.ptr(db), B.ptr(dx), 4, ctypes.byref(bst),
                                   None, None)
    assert code == B.E_ARG  # start flags are bits 0 and 1 only
