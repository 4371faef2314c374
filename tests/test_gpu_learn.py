"""Learned initial guess: BF16-MFMA MLP kernels against the NumPy oracle, and
the POD + MLP guess against the zero guess on a family of elasticity cells.

Parity bar for the network: forward outputs, loss and every gradient match
the oracle evaluated with the same BF16 rounding of the matrix-product
operands (float64 elsewhere) to 1e-3 relative in norm -- the remaining
difference is the FP32 accumulation and tanhf of the device. The oracle's
own backward pass is pinned by finite differences in test_oracle.py.
"""

import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.fixture(scope="module")
def net(B):
    nin, nh, nout = 4, 64, 40
    m = B.Mlp(nin, nh, nout, seed=3)
    rng = np.random.default_rng(7)
    _, npar = M.layout(nin, nh, nout)
    p = rng.normal(0, 0.3, npar).astype(np.float32)
    m.set_params(p)
    return m, p.astype(np.float64), (nin, nh, nout)


@pytest.mark.parametrize("batch", [1, 37, 256])
def test_mlp_forward_matches_oracle(B, ctx, net, batch):
    m, p, dims = net
    x = np.random.default_rng(batch).normal(size=(batch, dims[0])).astype(np.float32)
    y = m.forward(ctx, x)
    yo, _ = M.forward(p, x.astype(np.float64), *dims, bf16=True)
    assert relerr(y, yo) < 1e-3


def test_mlp_gradients_match_oracle(B, ctx, net):
    m, p, dims = net
    rng = np.random.default_rng(11)
    x = rng.normal(size=(83, dims[0])).astype(np.float32)
    t = rng.normal(size=(83, dims[2])).astype(np.float32)
    loss = m.backward(ctx, x, t)
    lo, go = M.loss_grad(p, x.astype(np.float64), t.astype(np.float64), *dims, bf16=True)
    g = m.grads().astype(np.float64)
    assert abs(loss - lo) <= 1e-4 * lo
    offs, _ = M.layout(*dims)
    for name, (o, shp) in zip(["W1", "b1", "W2", "b2", "W3", "b3"], offs):
        k = int(np.prod(shp))
        assert relerr(g[o:o + k], go[o:o + k]) < 1e-3, name


def test_mlp_adam_matches_oracle_and_learns(B, ctx):
    nin, nh, nout = 4, 32, 8
    m = B.Mlp(nin, nh, nout, seed=5)
    p0 = m.params().astype(np.float64)
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, size=(128, nin)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0] + k * x[:, 1]) * np.cos(x[:, 2] - x[:, 3] * k / 3) for k in range(nout)],
                 axis=1).astype(np.float32)
    m.backward(ctx, x, t)
    g = m.grads().astype(np.float64)
    m.adam_step(ctx, 1e-2)
    pn, _, _ = M.adam(p0, g, np.zeros_like(p0), np.zeros_like(p0), 1, 1e-2)
    assert relerr(m.params(), pn) < 1e-6
    first = m.backward(ctx, x, t)
    for _ in range(400):
        m.backward(ctx, x, t)
        m.adam_step(ctx, 3e-3)
    last = m.backward(ctx, x, t)
    print(f"mlp loss {first:.4f} -> {last:.5f}", flush=True)
    assert last < 0.1 * first


def test_learned_guess_beats_zero_guess(B, ctx):
    """A small family of 4^3 elasticity cells: the POD + MLP guess trained on
    solved cells lowers the PCG iteration count of unseen cells."""
    mgrid, ntrain, ntest = 4, 48, 8
    rng = np.random.default_rng(20261205)

    def draw(k):
        return np.stack([rng.uniform(0.1, 0.45, k), rng.uniform(0, 3, k), rng.uniform(0.2, 0.45, k),
                         rng.uniform(-0.15, 0.15, k)], axis=1)

    def cells(prm):
        out = []
        order = None
        for r, lc, nu, sh in prm:
            mat, rhs = B.gen_elastic_cell(mgrid, r, 10 ** lc, nu, sh, 20261205)
            order = mat.rcm_order() if order is None else order  # one template: one ordering
            mo = mat.reorder(order)
            out.append((mo, B.Factor.ic2(mo, B.default_params(tau1=0.1, tau2=0.01)), mo.reorder_vector(order, rhs)))
        return out

    def solve(cs, x0=None, fallback=False):
        bt = B.Batch([(mo, f) for mo, f, _ in cs], prec_mask=0b010)
        b = bt.gather([c[2] for c in cs])
        db = B.DeviceArray.from_host(b)
        dx = B.DeviceArray.from_host(x0) if x0 is not None else B.DeviceArray(b.shape)
        if x0 is None:
            dx.zero()
        st, its, rr = bt.solve(ctx, db, dx, B.default_opts(kind=B.BLOCKPCG, eps=1e-10, precision=B.F32), 6,
                               fallback)
        assert st.converged == st.pairs
        return dx, its

    ptrain, ptest = draw(ntrain), draw(ntest)
    train = cells(ptrain)
    n = train[0][0].n
    g = B.Guess(n, 6, 4, nmodes=16, nhidden=64, capacity=ntrain, seed=9)
    for k0 in range(0, ntrain, 16):
        dx, _ = solve(train[k0:k0 + 16])
        g.add(ctx, ptrain[k0:k0 + 16], dx)
    rep = g.fit(ctx, steps=1500, lr=3e-3)
    print(f"guess fit: modes {rep.modes} energy {rep.pod_energy:.6f} loss {rep.loss_first:.3e} -> "
          f"{rep.loss_final:.3e}", flush=True)
    assert rep.pod_energy > 0.9 and rep.loss_final < 0.2 * rep.loss_first
    test = cells(ptest)
    dx0, its0 = solve(test)
    x0 = B.DeviceArray((n * ntest, 6))
    g.predict(ctx, ptest, x0)
    guess = x0.host()
    exact = dx0.host()
    err = np.linalg.norm(guess - exact) / np.linalg.norm(exact)
    _, its1 = solve(test, x0=guess, fallback=True)
    print(f"guess rel err {err:.3e}; iterations zero {its0.mean():.1f} learned {its1.mean():.1f}", flush=True)
    assert err < 0.3
    assert its1.mean() < its0.mean()
