"""The CPU oracle, pinned before it is trusted.

* Against the reference's own compiled solve-phase code (oracle/_ref, built
  from Bslib's mvm.cpp, svector.cpp, cg.cpp, lanczos.cpp, rcm.cpp ...):
  bit-for-bit equality.
* Against the committed golden vectors (tests/golden/, written by the
  reference build) where the reference build is absent.
* Against closed forms: dense products and triangular solves, exact
  preconditioners (one iteration), direct solves.
"""

import os

import numpy as np
import pytest

from conftest import rel
from helpers import dense_factor, heat, seepage

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "solve_phase_small.npz")


def need_ref(O):
    if not O.ref_available():
        pytest.skip("reference build (oracle/_ref) not present")


@pytest.fixture(scope="module")
def hs(B, O):
    return heat(B, O)


@pytest.fixture(scope="module")
def ns(B, O):
    return seepage(B, O)


# ---- bit-exact against the reference build ----

def test_mvm_family_bitexact(O, hs, ns):
    need_ref(O)
    rng = np.random.default_rng(0)
    A = ns["A"]
    x = rng.standard_normal((A.n, 3))
    assert np.array_equal(O.mvma(A, x, 3), O.ref_mvma(A, x, 3))
    assert np.array_equal(O.mvmat(A, x, 3), O.ref_mvmat(A, x, 3))
    S = hs["A"]
    xs = rng.standard_normal(S.n)
    assert np.array_equal(O.mvmasymm(S, xs), O.ref_mvmasymm(S, xs))


def test_triangular_bitexact(O, hs, ns):
    need_ref(O)
    rng = np.random.default_rng(1)
    for f in (hs["U"], ns["L"], ns["U"]):
        x = rng.standard_normal((f.n, 2))
        assert np.array_equal(O.solvel(f, x, 2), O.ref_solvel(f, x, 2))
        assert np.array_equal(O.solveu(f, x, 2), O.ref_solveu(f, x, 2))


@pytest.mark.parametrize("ittype,eps", [(0, 1e-10), (1, 1e-6), (2, 1e-7)])
def test_cg_bitexact(O, hs, ittype, eps):
    need_ref(O)
    x0, r = O.cg(hs["A"], hs["U"], hs["b"], eps=eps, ittype=ittype)
    x1, nit, _ = O.ref_cg(hs["A"], hs["U"], hs["b"], eps=eps, ittype=ittype)
    assert r.nitperf == nit
    assert np.array_equal(x0, x1)


@pytest.mark.parametrize("sttype,eps", [(0, 1e-8), (1, 1e-10), (2, 1e-9)])
def test_lanczos_bitexact(O, ns, sttype, eps):
    need_ref(O)
    x0, r = O.lanczos(ns["A"], ns["L"], ns["U"], ns["b"], eps=eps, sttype=sttype)
    x1, nit, _ = O.ref_lanczos(ns["A"], ns["L"], ns["U"], ns["b"], eps=eps, sttype=sttype)
    assert r.nitperf == nit
    assert np.array_equal(x0, x1)


def test_scprod_bitexact(O):
    need_ref(O)
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(1001), rng.standard_normal(1001)
    assert O.orc().orc_scprod(1001, O._dp(x), O._dp(y)) == O.ref().ref_scprod(1001, O._dp(x), O._dp(y))


# ---- golden vectors (written by the reference build) ----

def test_golden_vectors(O):
    if not os.path.exists(GOLDEN):
        pytest.skip("golden fixture not generated")
    g = np.load(GOLDEN, allow_pickle=False)
    A = O.Block(g["a_sprnd"], g["a_ia"], g["a_ja"], g["a_a"])
    U = O.Block(g["a_sprnd"], g["u_ia"], g["u_ja"], g["u_a"])
    assert np.array_equal(O.mvmasymm(A, g["x"]), g["mvmasymm"])
    assert np.array_equal(O.solvel(U, g["x"]), g["solvel"])
    assert np.array_equal(O.solveu(U, g["x"]), g["solveu"])
    x, r = O.cg(A, U, g["b"], eps=1e-10, ittype=0)
    assert r.nitperf == int(g["cg_nit"])
    assert np.array_equal(x, g["cg_x"])
    N = O.Block(g["n_sprnd"], g["n_ia"], g["n_ja"], g["n_a"])
    L = O.Block(g["n_sprnd"], g["l_ia"], g["l_ja"], g["l_a"])
    Un = O.Block(g["n_sprnd"], g["nu_ia"], g["nu_ja"], g["nu_a"])
    assert np.array_equal(O.mvma(N, g["xn"]), g["mvma"])
    assert np.array_equal(O.mvmat(N, g["xn"]), g["mvmat"])
    xl, rl = O.lanczos(N, L, Un, g["bn"], eps=1e-10, sttype=1)
    assert rl.nitperf == int(g["lz_nit"])
    assert np.array_equal(xl, g["lz_x"])


# ---- closed forms ----

def test_products_closed_form(B, O, hs, ns):
    rng = np.random.default_rng(3)
    Ad = ns["mo"].dense()
    x = rng.standard_normal(Ad.shape[0])
    assert rel(O.mvma(ns["A"], x), Ad @ x) < 1e-14
    assert rel(O.mvmat(ns["A"], x), Ad.T @ x) < 1e-14
    Sd = hs["mo"].dense()
    xs = rng.standard_normal(Sd.shape[0])
    assert rel(O.mvmasymm(hs["A"], xs), Sd @ xs) < 1e-14


def test_triangular_closed_form(O, hs, ns):
    rng = np.random.default_rng(4)
    for f in (hs["U"], ns["L"], ns["U"]):
        D = dense_factor(f)
        x = rng.standard_normal(f.n)
        assert rel(O.solvel(f, x), np.linalg.solve(D.T, x)) < 1e-12
        assert rel(O.solveu(f, x), np.linalg.solve(D, x)) < 1e-12


def test_exact_preconditioner_one_step(B, O):
    """tau1 = tau2 = theta = 0 keeps every block: the factors are exact and every
    Krylov method stops after one correction."""
    h = heat(B, O, 5, 4, 6, seed=9, tau1=0.0, tau2=0.0, theta=0.0)
    x, r = O.cg(h["A"], h["U"], h["b"], eps=1e-10, ittype=0)
    assert r.updates == 1
    assert rel(x, np.linalg.solve(h["mo"].dense(), h["b"])) < 1e-12
    s = seepage(B, O, 4, 3, 3, 4, seed=5, tau1=0.0, tau2=0.0, theta=0.0)
    xd = np.linalg.solve(s["mo"].dense(), s["b"])
    xb, rb = O.bicgstab(s["A"], s["L"], s["U"], s["b"], eps=1e-10)
    assert rb.nitperf == 1 and rel(xb, xd) < 1e-12
    xg, rg = O.fgmres(s["A"], s["L"], s["U"], s["b"], eps=1e-10)
    assert rg.nitperf == 1 and rel(xg, xd) < 1e-12


def test_krylov_against_direct(O, hs, ns):
    xd = np.linalg.solve(hs["mo"].dense(), hs["b"])
    x, r = O.cg(hs["A"], hs["U"], hs["b"], eps=1e-12, ittype=3)
    assert rel(x, xd) < 1e-9
    Nd = ns["mo"].dense()
    xd = np.linalg.solve(Nd, ns["b"])
    for fn in (O.bicgstab, O.fgmres):
        x, r = fn(ns["A"], ns["L"], ns["U"], ns["b"], eps=1e-12)
        assert np.linalg.norm(ns["b"] - Nd @ x) <= 1.01e-12 * np.linalg.norm(ns["b"])
        assert rel(x, xd) < 1e-9
    x, r = O.lanczos(ns["A"], ns["L"], ns["U"], ns["b"], eps=1e-12, sttype=1)
    assert rel(x, xd) < 1e-8


def test_mlp_oracle_gradient_matches_finite_differences():
    """Pins the learned-guess network's backward pass (the GPU gradients'
    checker) against central differences of its own float64 loss."""
    import mlp_oracle as M

    rng = np.random.default_rng(5)
    nin, nh, nout, batch = 4, 7, 5, 6
    _, npar = M.layout(nin, nh, nout)
    p = rng.normal(0, 0.5, npar)
    x = rng.normal(size=(batch, nin))
    t = rng.normal(size=(batch, nout))
    loss, g = M.loss_grad(p, x, t, nin, nh, nout)
    h = 1e-6
    num = np.empty(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        num[i] = (M.loss_grad(p + e, x, t, nin, nh, nout)[0] - M.loss_grad(p - e, x, t, nin, nh, nout)[0]) / (2 * h)
    assert np.max(np.abs(num - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # the loss is the closed form of the network
    w1, b1, w2, b2, w3, b3 = M.unpack(p, nin, nh, nout)
    assert abs(loss - np.mean((np.tanh(np.tanh(x @ w1 + b1) @ w2 + b2) @ w3 + b3 - t) ** 2)) < 1e-14


def test_mlp_oracle_bf16_rounding():
    import mlp_oracle as M

    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -3.14159, 1e-30, 65504.0])
    r = M.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -6  # 1 + 3/256 rounds to 1 + 4/256
    assert abs(r[3] + 3.140625) < 1e-12
    assert np.all(np.abs(r - v) <= np.abs(v) * 2 ** -8)


def test_cg_par_oracle_pins(B, O):
    """Parallel-form CG (cg.cpp:355-771): with an exact factor it stops after
    one step, and the reference's exit before the last solution step leaves
    x0 untouched (apply_last=False) while the applied step gives the solution;
    on a general system it follows the serial Cg restatement (same iterates in
    exact arithmetic)."""
    h = heat(B, O, 6, 5, 9, seed=41, tau1=0.0, tau2=0.0, theta=0.0)  # exact block Cholesky
    x1, r1 = O.cg_par(h["A"], h["U"], h["b"], eps=1e-12, sttype=1, apply_last=True)
    x0, r0 = O.cg_par(h["A"], h["U"], h["b"], eps=1e-12, sttype=1, apply_last=False)
    xd = np.linalg.solve(h["mo"].dense(), h["b"])
    assert r1.nitperf == 0 and r1.updates == 1
    assert np.linalg.norm(x1 - xd) <= 1e-10 * np.linalg.norm(xd)
    assert np.all(x0 == 0.0)
    g = heat(B, O, 12, 10, 9, seed=42)
    xp, rp = O.cg_par(g["A"], g["U"], g["b"], eps=1e-11, sttype=1, apply_last=True)
    xs, rs = O.cg(g["A"], g["U"], g["b"], eps=1e-11, ittype=0)
    assert abs(rp.updates - rs.updates) <= 1
    assert np.linalg.norm(xp - xs) <= 1e-8 * np.linalg.norm(xs)


def _block_gmres_brute(Ad, Minv, R0, k):
    """Minimal-residual correction over the block Krylov space
    span{R0, (A M^-1) R0, ...} of k blocks, built by block Arnoldi with full
    reorthogonalisation; returns (correction, per-column residual norms)."""
    AM = Ad @ Minv
    Q, _ = np.linalg.qr(R0)
    basis = [Q]
    for _ in range(k - 1):
        W = AM @ basis[-1]
        for _ in range(2):
            for V in basis:
                W = W - V @ (V.T @ W)
        Q, _ = np.linalg.qr(W)
        basis.append(Q)
    Vk = np.hstack(basis)
    Z, *_ = np.linalg.lstsq(AM @ Vk, R0, rcond=None)
    return Minv @ (Vk @ Z), np.linalg.norm(R0 - AM @ Vk @ Z, axis=0)


def test_block_gmres_oracle_pins(B, O):
    """Block GMRES (gmres.cpp:337-1038), pinned by what the mathematics fixes:
    after k steps each column is the minimal-residual correction over the
    block Krylov space of k blocks (brute force by dense least squares), the
    recursion's residual norms are those minimal residuals, the exit step is
    the first one at which every column passes its test, and an exact factor
    converges in one step."""
    s = seepage(B, O, seed=23)
    n = s["A"].n
    Ad = s["mo"].dense()
    I = np.eye(n)
    Minv = O.solveu(s["U"], O.solvel(s["L"], I, n), n)
    rng = np.random.default_rng(5)
    Bm = rng.standard_normal((n, 3))
    X0 = 0.1 * rng.standard_normal((n, 3))
    R0 = Bm - Ad @ X0
    for k in (1, 2, 4):
        x, r, hist, fin = O.block_gmres(s["A"], s["L"], s["U"], Bm, X0, niter=k, eps=0.0)
        corr, res = _block_gmres_brute(Ad, Minv, R0, k)
        assert r.nitperf == k and r.updates == 0
        assert np.linalg.norm(x - X0 - corr) <= 1e-10 * np.linalg.norm(corr), k
        assert np.allclose(hist[k - 1], res, rtol=1e-8, atol=0), (k, hist[k - 1], res)
        assert np.allclose(fin, np.linalg.norm(Bm - Ad @ x, axis=0), rtol=1e-12)
        assert np.allclose(fin, res, rtol=1e-6)
    # exit step: relative (sttype 1) and absolute (sttype 0) tests
    x, r, hist, fin = O.block_gmres(s["A"], s["L"], s["U"], Bm, X0, niter=60, eps=1e-9, sttype=1)
    r0n = np.linalg.norm(R0, axis=0)
    first = next(k for k in range(60) if np.all(hist[k] / r0n < 1e-9))
    assert r.updates == 1 and r.nitperf == first + 1
    assert np.all(fin <= 1e-8 * r0n)
    xa, ra, ha, fa = O.block_gmres(s["A"], s["L"], s["U"], Bm, X0, niter=60, eps=1e-7, sttype=0)
    first = next(k for k in range(60) if np.all(ha[k] < 1e-7))
    assert ra.updates == 1 and ra.nitperf == first + 1
    # single column: the block method is plain right-preconditioned GMRES
    x1, r1, h1, f1 = O.block_gmres(s["A"], s["L"], s["U"], Bm[:, :1], X0[:, :1], niter=3, eps=0.0)
    c1, res1 = _block_gmres_brute(Ad, Minv, R0[:, :1], 3)
    assert np.linalg.norm(x1 - X0[:, :1] - c1) <= 1e-10 * np.linalg.norm(c1)
    # exact factor: one step, the direct solution
    e = seepage(B, O, 4, 3, 3, 4, seed=5, tau1=0.0, tau2=0.0, theta=0.0)
    Be = rng.standard_normal((e["A"].n, 4))
    xe, re, he, fe = O.block_gmres(e["A"], e["L"], e["U"], Be, niter=10, eps=1e-10)
    assert re.nitperf == 1 and re.updates == 1
    assert np.linalg.norm(xe - np.linalg.solve(e["mo"].dense(), Be)) <= 1e-11 * np.linalg.norm(xe)


# ---- pins of the restatements that have no reference counterpart ----

def test_mlp_oracle_adam_matches_torch():
    """mlp_oracle.adam (the rule the GPU Adam kernel is checked against)
    against torch.optim.Adam in float64 on the CPU, over a sequence of steps
    with changing gradients (bias correction by step count)."""
    torch = pytest.importorskip("torch")
    import mlp_oracle as M

    rng = np.random.default_rng(7)
    p = rng.standard_normal(257)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 26):
        g = rng.standard_normal(p.size) * (1.0 + step)
        p, m, v = M.adam(p, g, m, v, step, 3e-3)
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_bicgstab_oracle_iterations_match_scipy(B, O, ns):
    """orc_bicgstab (right-preconditioned BiCGStab, the nonsymmetric drop-in's
    method) against scipy.sparse.linalg.bicgstab with the same explicit
    preconditioner M^-1 = U^-1 L^-1 and the same absolute stopping level
    eps ||b||: iteration counts within one (scipy counts a half-step exit as
    an iteration the same way only when the callback sees it) and the same
    solution to the tolerance."""
    spla = pytest.importorskip("scipy.sparse.linalg")
    A = ns["mo"].dense()
    minv = lambda r: O.solveu(ns["U"], O.solvel(ns["L"], r))   # noqa: E731
    M = spla.LinearOperator(A.shape, matvec=minv, dtype=np.float64)
    b = ns["b"]
    for eps in (1e-8, 1e-11):
        count = [0]
        xs, info = spla.bicgstab(A, b, rtol=0.0, atol=eps * np.linalg.norm(b), M=M, maxiter=500,
                                 callback=lambda xk: count.__setitem__(0, count[0] + 1))
        assert info == 0
        xo, r = O.bicgstab(ns["A"], ns["L"], ns["U"], b, eps=eps)
        assert abs(r.nitperf - count[0]) <= 1, (eps, r.nitperf, count[0])
        assert rel(xo, xs) < 1e3 * eps


def test_fgmres_oracle_is_the_minimal_residual_iterate(B, O, ns):
    """orc_fgmres with a fixed preconditioner is right-preconditioned GMRES:
    after k < m steps its iterate minimises ||b - A x|| over x0 + M^-1
    K_k(A M^-1, r0). Checked against the dense least-squares minimiser over
    that space for k = 1..8 (a dropped Hessenberg term, a wrong rotation sign
    or a left-preconditioned basis fail it)."""
    A = ns["mo"].dense()
    n = A.shape[0]
    minv = lambda r: O.solveu(ns["U"], O.solvel(ns["L"], r))   # noqa: E731
    b = ns["b"]
    x0 = np.random.default_rng(3).standard_normal(n) * 0.1
    r0 = b - A @ x0
    basis = [r0 / np.linalg.norm(r0)]
    for k in range(1, 9):
        Q, _ = np.linalg.qr(np.stack(basis, axis=1))
        Z = np.stack([minv(Q[:, j]) for j in range(k)], axis=1)
        y, *_ = np.linalg.lstsq(A @ Z, r0, rcond=None)
        xbf = x0 + Z @ y
        xo, r = O.fgmres(ns["A"], ns["L"], ns["U"], b, x0=x0, m=30, niter=k, eps=1e-300)
        assert rel(xo, xbf) < 1e-9, k
        basis.append(A @ minv(Q[:, k - 1]))


@pytest.mark.parametrize("variant", [0, 1, 2], ids=["right", "left", "central"])
@pytest.mark.parametrize("nrhs", [1, 3])
def test_block_lanczos_oracle_matches_reference(B, O, variant, nrhs):
    """orc_block_lanczos against the reference's own BlockLanczosRight / Left /
    Central (lanczos.cpp:932-3225, compiled with its LAPACK into oracle/_ref)
    on the same ILU2-preconditioned system: the same solution to 1e-12 (the
    restatement inverts rho by elimination where the reference uses an SVD,
    so the iterates agree to rounding, not bit for bit), at a loose and a
    tight tolerance; and the solution solves the system."""
    need_ref(O)
    s = seepage(B, O, 5, 4, 3, 6, seed=21)
    b = np.random.default_rng(5 + nrhs).standard_normal((s["A"].n, nrhs))
    A = s["mo"].dense()
    for eps in (1e-6, 1e-11):
        xr = O.ref_block_lanczos(variant, s["A"], s["L"], s["U"], b, eps=eps)
        xo, r = O.block_lanczos(variant, s["A"], s["L"], s["U"], b, eps=eps)
        assert r.updates == 1 and r.breakdown == 0
        assert rel(xo, xr) < 1e-12, (variant, eps)
        res = max(np.linalg.norm(b[:, c] - A @ xo[:, c]) / np.linalg.norm(b[:, c]) for c in range(nrhs))
        assert res < 10 * eps
