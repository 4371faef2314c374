"""Host stages of the path: assembly, RCM ordering, reordering, L/U split and
the IC2 / ILU2 factorizations.

Ordering, structural transforms and the factorizations are compared bit for
bit with the reference's own compiled code (oracle/_ref: the Bslib sources
plus the reference's LAPACK 3.0, lapacksm/*.f and lapackdm.f, compiled with
flang). The IC2 / ILU2 factors must equal the reference's Ich2 / Ilu2 in
pattern and in every value, and the dense DSYEV / DGESVD restatements must
equal the reference's LAPACK on adversarial blocks. Independent pins remain
beside them: the special case that reduces the factorizations to exact block
Cholesky / LU (tau1 = tau2 = theta = 0 keeps every block), and a dense numpy
restatement of the second-order algorithm compared through the
preconditioner operator.
"""

import numpy as np
import pytest

from conftest import rel
from helpers import dense_factor, heat, seepage


def need_ref(O):
    if not O.ref_available():
        pytest.skip("reference build (oracle/_ref) not present")


# ---- generators ----

def test_generator_shapes_and_determinism(B):
    m, r = B.gen_heat2d(64, 64, 9, 20261202)
    i = m.info()
    assert (i["nsup"], i["n"], i["nzja"]) == (4096, 36864, 12160)   # config-0 shape, upper storage
    m2, r2 = B.gen_heat2d(64, 64, 9, 20261202)
    assert np.array_equal(m.arrays()[3], m2.arrays()[3]) and np.array_equal(r, r2)
    s, _ = B.gen_seepage3d(6, 5, 4, 3, 0.3, 1)
    assert s.info()["nzja"] == 120 + 2 * (5 * 5 * 4 + 6 * 4 * 4 + 6 * 5 * 3)   # 7-point, non-periodic
    e, r6 = B.gen_elastic_cell(4, 0.3, 10.0, 0.3, 0.0, 7)
    assert e.info()["nzja"] == 64 * 4 and r6.shape == (64 * 24, 6)       # periodic: diag + 3 upper


def test_generators_spd_and_skew(B):
    m, _ = B.gen_heat2d(5, 4, 4, 2)
    d = m.dense()
    assert np.allclose(d, d.T) and np.linalg.eigvalsh(d).min() > 0
    s, _ = B.gen_seepage3d(3, 3, 2, 4, 0.3, 2)
    ds = s.dense()
    assert np.linalg.norm(ds - ds.T) > 0.1 * np.linalg.norm(ds + ds.T) / np.sqrt(ds.shape[0])
    e, _ = B.gen_elastic_cell(3, 0.3, 100.0, 0.4, 0.0, 9)
    de = e.dense()
    assert np.allclose(de, de.T) and np.linalg.eigvalsh(de).min() > 0


# ---- ordering and transforms against the reference build ----

@pytest.mark.parametrize("kind", ["heat", "seepage", "elastic", "blockdiag"])
def test_rcm_matches_reference(B, O, kind):
    need_ref(O)
    if kind == "heat":
        m, _ = B.gen_heat2d(13, 9, 2, 4)
    elif kind == "seepage":
        m, _ = B.gen_seepage3d(6, 5, 4, 2, 0.3, 4)
    elif kind == "elastic":
        m, _ = B.gen_elastic_cell(5, 0.3, 10.0, 0.3, 0.0, 4)
    else:  # two disconnected chains plus an isolated block row
        sprnd = np.arange(0, 2 * 9, 2)
        ia = [0, 2, 4, 6, 7, 9, 11, 12, 13]
        ja = [0, 1, 1, 2, 2, 3, 3, 4, 5, 5, 6, 6, 7]
        m = B.Matrix.from_arrays(sprnd, ia, ja, np.ones(4 * 13), symmetric=True)
    A = O.Block(*m.arrays())
    assert np.array_equal(m.rcm_order(), O.ref_order_prf(A))


def test_genrcm_small_graphs(O, B):
    """Path, star and grid graphs (self loops kept, as OrderPrfMtr feeds genrcm)."""
    need_ref(O)
    for edges, n in (([(i, i + 1) for i in range(6)], 7), ([(0, i) for i in range(1, 6)], 6),
                     ([(r * 4 + c, r * 4 + c + 1) for r in range(3) for c in range(3)]
                      + [(r * 4 + c, r * 4 + c + 4) for r in range(2) for c in range(4)], 12)):
        adj = [{i} for i in range(n)]
        for u, v in edges:
            adj[u].add(v); adj[v].add(u)
        ia = np.concatenate([[0], np.cumsum([len(s) for s in adj])]).astype(np.int32)
        ja = np.concatenate([sorted(s) for s in adj]).astype(np.int32)
        upper_ja = [[c for c in sorted(s) if c >= i] for i, s in enumerate(adj)]
        uia = np.concatenate([[0], np.cumsum([len(s) for s in upper_ja])]).astype(np.int32)
        m = B.Matrix.from_arrays(np.arange(n + 1), uia, np.concatenate(upper_ja), np.ones(uia[-1]), True)
        perm = O.ref_genrcm(n, ia, ja)
        order = np.empty(n, np.int32)
        order[perm] = np.arange(n)
        assert np.array_equal(m.rcm_order(), order)


def test_reorder_and_combine_lu_match_reference(B, O):
    need_ref(O)
    h = heat(B, O, 7, 6, 3, seed=5)
    A = O.Block(*h["m"].arrays())
    ia, ja, a = O.ref_ordmtr(A, h["order"], True)
    assert np.array_equal(ia, h["A"].ia) and np.array_equal(ja, h["A"].ja) and np.array_equal(a, h["A"].a)
    s = seepage(B, O, 4, 4, 3, 3, seed=6)
    N = O.Block(*s["m"].arrays())
    ia, ja, a = O.ref_ordmtr(N, s["order"], False)
    assert np.array_equal(ja, s["A"].ja) and np.array_equal(a, s["A"].a)
    # CombineLU: the factor's input patterns (L by columns, U by rows, union-padded)
    li, lj, la = O.ref_combinelu(s["A"], 0)
    ui, uj, ua = O.ref_combinelu(s["A"], 1)
    assert np.array_equal(li, ui) and np.array_equal(lj, uj)
    assert lj.size >= s["A"].nsup


@pytest.mark.parametrize("kind", ["seepage-b3", "seepage-b24", "seepage-skew0"])
def test_combine_lu_matches_reference(B, O, kind):
    """The product's combine_lu (through bcm_combine_lu) against the
    reference's own CombineLU (smatrixrs.cpp:1645-1803) built in oracle/_ref:
    patterns and values bit for bit, for both halves."""
    need_ref(O)
    bla, skew, dims = {"seepage-b3": (3, 0.3, (4, 4, 3)), "seepage-b24": (24, 0.3, (3, 3, 2)),
                       "seepage-skew0": (6, 0.0, (5, 3, 2))}[kind]
    s = seepage(B, O, *dims, bla, skew, seed=17)
    lo, up = s["mo"].combine_lu()
    for which, mine in ((0, lo), (1, up)):
        ri, rj, ra = O.ref_combinelu(s["A"], which)
        sp, ia, ja, a = mine.arrays()
        assert np.array_equal(sp, s["A"].sprnd)
        assert np.array_equal(ia, ri) and np.array_equal(ja, rj), which
        assert np.array_equal(a, ra), which


def test_vector_reorder_roundtrip(B):
    m, rhs = B.gen_heat2d(6, 6, 3, 8)
    o = m.rcm_order()
    v = m.reorder_vector(o, rhs)
    assert not np.array_equal(v, rhs)
    assert np.array_equal(m.reorder_vector(o, v, inverse=True), rhs)


# ---- factorizations ----

def test_ic2_exact_special_case(B, O):
    h = heat(B, O, 6, 5, 5, seed=7, tau1=0.0, tau2=0.0, theta=0.0)
    U = dense_factor(h["U"])
    A = h["mo"].dense()
    assert rel(U.T @ U, A) < 1e-13


def test_ilu2_exact_special_case(B, O):
    s = seepage(B, O, 4, 3, 3, 5, seed=8, tau1=0.0, tau2=0.0, theta=0.0)
    L = dense_factor(s["L"]).T
    U = dense_factor(s["U"])
    A = s["mo"].dense()
    assert rel(L @ U, A) < 1e-13


def _blocks(Bk):
    sp = Bk.sprnd
    out = {}
    off = 0
    for i in range(Bk.nsup):
        bi = sp[i + 1] - sp[i]
        for j in range(Bk.ia[i], Bk.ia[i + 1]):
            c = Bk.ja[j]
            bj = sp[c + 1] - sp[c]
            out[(i, c)] = Bk.a[off:off + bi * bj].reshape(bj, bi).T.copy()
            off += bi * bj
    return out


def _numpy_ic2(A, tau1, tau2, theta):
    """Dense-block restatement of IC2 (scaled space, second-order updates)."""
    blk = _blocks(A)
    n = A.nsup
    S, Sinv = {}, {}
    for i in range(n):
        w, Q = np.linalg.eigh(blk[(i, i)])
        S[i] = np.diag(1 / np.sqrt(w)) @ Q.T
        Sinv[i] = Q @ np.diag(np.sqrt(w))
    Dg = {i: S[i] @ blk[(i, i)] @ S[i].T for i in range(n)}
    G = {}
    for i in range(n):
        F = {i: Dg[i].copy()}
        for (r, c), b in blk.items():
            if r == i and c > i:
                F[c] = S[i] @ b @ S[c].T
        for p in range(i):
            if (p, i) not in G:
                continue
            lpi, rpi = G[(p, i)]
            for (pp, j), (gpj, rpj) in list(G.items()):
                if pp != p or j < i or (rpi and rpj):
                    continue
                F[j] = F.get(j, 0) - lpi.T @ gpj
        for j in sorted(k for k in F if k != i):
            if np.abs(F[j]).max() < tau2:
                v = theta * np.abs(F[j])
                F[i] += np.diag(v.sum(axis=1))
                Dg[j] += np.diag(v.sum(axis=0))
                del F[j]
        w, Q = np.linalg.eigh(F[i])
        delem = np.diag(1 / np.sqrt(w)) @ Q.T
        G[(i, i)] = (np.linalg.inv(delem.T), False)
        for j in sorted(k for k in F if k != i):
            g = delem @ F[j]
            G[(i, j)] = (g, np.abs(g).max() < tau1)
    U = np.zeros((A.n, A.n))
    sp = A.sprnd
    pattern = set()
    for (i, j), (g, isr) in G.items():
        if isr:
            continue
        pattern.add((i, j))
        U[sp[i]:sp[i + 1], sp[j]:sp[j + 1]] = g @ Sinv[j].T
    return U, pattern


def test_ic2_against_numpy_restatement(B, O):
    h = heat(B, O, 4, 3, 3, seed=12, tau1=3e-2, tau2=3e-3, theta=0.2)
    Un, pattern = _numpy_ic2(h["A"], 3e-2, 3e-3, 0.2)
    Ud = dense_factor(h["U"])
    assert pattern == set(_blocks(h["U"]).keys())
    assert rel(Ud.T @ Ud, Un.T @ Un) < 1e-12
    # not the exact factor: the drop rules were exercised
    assert rel(Ud.T @ Ud, h["mo"].dense()) > 1e-6


def _numpy_ilu2(A, tau1, tau2, theta, sclmin=1e-8, pivmin=1e-8):
    blk = _blocks(A)
    n = A.nsup
    scll, sclu, sinvl, sinvu = {}, {}, {}, {}
    for i in range(n):
        u, s, vt = np.linalg.svd(blk[(i, i)])
        s2 = np.sqrt(np.maximum(s, sclmin))
        scll[i] = np.diag(1 / s2) @ u.T
        sclu[i] = np.diag(1 / s2) @ vt
        sinvl[i] = u @ np.diag(s2)
        sinvu[i] = vt.T @ np.diag(s2)
    Dg = {i: scll[i] @ blk[(i, i)] @ sclu[i].T for i in range(n)}
    GL, GU, R = {}, {}, {}
    for i in range(n):
        FU, FL = {i: Dg[i].copy()}, {i: Dg[i].copy()}
        for (r, c), b in blk.items():
            if r == i and c > i:
                FU[c] = scll[i] @ b @ sclu[c].T
            if r > i and c == i:   # lower block (r, i) enters row i transposed
                FL[r] = sclu[i] @ b.T @ scll[r].T
        for p in range(i):
            if (p, i) not in GU:
                continue
            lpi, upi, rpi = GL[(p, i)], GU[(p, i)], R[(p, i)]
            for (pp, j) in list(GU.keys()):
                if pp != p or j < i or (rpi and R[(p, j)]):
                    continue
                FU[j] = FU.get(j, 0) - lpi.T @ GU[(p, j)]
                FL[j] = FL.get(j, 0) - upi.T @ GL[(p, j)]
        for j in sorted(k for k in FU if k != i):
            FL.setdefault(j, np.zeros_like(FU[j]))
            if max(np.abs(FL[j]).max(), np.abs(FU[j]).max()) < tau2:
                v = theta * np.maximum(np.abs(FL[j]), np.abs(FU[j]))
                FU[i] += np.diag(v.sum(axis=1))
                Dg[j] += np.diag(v.sum(axis=0))
                del FU[j], FL[j]
        u, s, vt = np.linalg.svd(FU[i])
        e = 1 / np.sqrt(np.maximum(s, pivmin))
        dl, du = np.diag(e) @ u.T, np.diag(e) @ vt
        GL[(i, i)], GU[(i, i)], R[(i, i)] = dl.T, du.T, False
        for j in sorted(k for k in FU if k != i):
            GL[(i, j)], GU[(i, j)] = du @ FL[j], dl @ FU[j]
            R[(i, j)] = max(np.abs(GL[(i, j)]).max(), np.abs(GU[(i, j)]).max()) < tau1
    L = np.zeros((A.n, A.n))
    U = np.zeros((A.n, A.n))
    sp = A.sprnd
    pattern = set()
    for (i, j) in GU:
        if R[(i, j)]:
            continue
        pattern.add((i, j))
        if i == j:
            L[sp[i]:sp[i + 1], sp[i]:sp[i + 1]] = np.linalg.inv(scll[i].T @ GL[(i, i)]).T
            U[sp[i]:sp[i + 1], sp[i]:sp[i + 1]] = np.linalg.inv(sclu[i].T @ GU[(i, i)])
        else:
            L[sp[j]:sp[j + 1], sp[i]:sp[i + 1]] = (GL[(i, j)] @ sinvl[j].T).T
            U[sp[i]:sp[i + 1], sp[j]:sp[j + 1]] = GU[(i, j)] @ sinvu[j].T
    return L, U, pattern


def test_ilu2_against_numpy_restatement(B, O):
    s = seepage(B, O, 3, 3, 2, 3, seed=13, tau1=3e-2, tau2=3e-3, theta=0.2)
    Ln, Un, pattern = _numpy_ilu2(s["A"], 3e-2, 3e-3, 0.2)
    L = dense_factor(s["L"]).T
    U = dense_factor(s["U"])
    assert pattern == set(_blocks(s["U"]).keys())
    assert rel(L @ U, Ln @ Un) < 1e-12
    assert rel(L @ U, s["mo"].dense()) > 1e-6


def test_ic2_rejects_indefinite(B):
    sprnd = np.array([0, 2, 4])
    ia = np.array([0, 2, 3])
    ja = np.array([0, 1, 1])
    a = np.array([1.0, 0, 0, -1.0, 0.1, 0, 0, 0.1, 1.0, 0, 0, 1.0])
    m = B.Matrix.from_arrays(sprnd, ia, ja, a, symmetric=True)
    with pytest.raises(B.BcmError) as e:
        B.Factor.ic2(m)
    assert e.value.code == B.E_NOT_SPD


def test_ic2_workspace_bound(B):
    m, _ = B.gen_heat2d(8, 8, 3, 1)
    mo = m.reorder(m.rcm_order())
    with pytest.raises(B.BcmError) as e:
        B.Factor.ic2(mo, B.default_params(memory=0.3))
    assert e.value.code == B.E_WORKSPACE
    f = B.Factor.ic2(mo, B.default_params(memory=4.0))   # bounded but sufficient: same factor
    g = B.Factor.ic2(mo)
    assert np.array_equal(f.arrays(1, 64)[2], g.arrays(1, 64)[2])


# ---- dense kernels and factors against the reference's own LAPACK 3.0 build ----

def _dense_cases(n, rng):
    """Matrices that stress the eigen/singular solvers: random, repeated
    eigenvalues in a random basis (isotropic blocks), exactly diagonal with
    ties, rank deficient, graded, and scaled into DLASCL's ranges."""
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    g = rng.standard_normal((n, n))
    rep = np.repeat(np.arange(1, (n + 2) // 3 + 1, dtype=float), 3)[:n]
    cases = {
        "random": g,
        "repeated": (q * rep) @ q.T,
        "diag-ties": np.diag(np.repeat([3.0, 1.0], (n + 1) // 2)[:n]),
        "rank-def": g[:, : max(1, n // 2)] @ g[: max(1, n // 2), :],
        "graded": g * np.logspace(0, -12, n)[None, :],
        "tiny": g * 1e-160,
        "huge": g * 1e160,
        "zero": np.zeros((n, n)),
        "identity": np.eye(n),
    }
    return cases


@pytest.mark.parametrize("n", [1, 2, 3, 9, 24, 32])
def test_dsyev_matches_reference_lapack(B, O, n):
    """bcm_dsyev (the product's LAPACK 3.0 DSYEV restatement) against
    dsyev_("V", "U") of the reference (lapackd.cpp -> lapackdm.f ->
    lapacksm/dsyev.f, built with flang): eigenvalues and eigenvectors bit for
    bit, including repeated eigenvalues, where the basis is arbitrary."""
    need_ref(O)
    rng = np.random.default_rng(100 + n)
    for name, a in _dense_cases(n, rng).items():
        s = a + a.T
        w, z = B.dsyev(s)
        wr, zr, info = O.ref_dsyev(s)
        assert info == 0, name
        assert np.array_equal(w, wr) and np.array_equal(z, zr), (n, name)
        if name == "random":  # and it is an eigen-decomposition
            assert np.allclose(z @ np.diag(w) @ z.T, s, atol=1e-12 * np.abs(s).max())


@pytest.mark.parametrize("n", [1, 2, 3, 9, 24, 32])
def test_dgesvd_matches_reference_lapack(B, O, n):
    """bcm_dgesvd against dgesvd_("A", "A") of the reference's LAPACK 3.0:
    singular values and both singular vector sets bit for bit."""
    need_ref(O)
    rng = np.random.default_rng(200 + n)
    for name, a in _dense_cases(n, rng).items():
        s, u, vt = B.dgesvd(a)
        sr, ur, vtr, info = O.ref_dgesvd(a)
        assert info == 0, name
        assert np.array_equal(s, sr) and np.array_equal(u, ur) and np.array_equal(vt, vtr), (n, name)
        if name == "random":
            assert np.allclose((u * s) @ vt, a, atol=1e-12 * np.abs(a).max())


def _fct_params(B, tau):
    tau1, tau2 = tau
    return B.default_params(tau1=tau1, tau2=tau2, memory=10.0)


@pytest.mark.parametrize("kind,tau", [("heat", (1e-2, 1e-4)), ("heat", (0.1, 0.01)),
                                      ("elastic", (1e-2, 1e-4)), ("elastic", (0.1, 0.01))])
def test_ic2_bitwise_equals_reference(B, O, kind, tau):
    """The product's IC2 factor (pattern and every value) equals the
    reference's own CSMatrixRS::Ich2 (ich2rs.cpp:1075-1378) compiled with its
    LAPACK 3.0, on the same RCM-ordered upper-stored matrix."""
    need_ref(O)
    if kind == "heat":
        m, _ = B.gen_heat2d(10, 9, 9, 41)
    else:
        m, _ = B.gen_elastic_cell(4, 0.35, 300.0, 0.3, 0.05, 42)
    mo = m.reorder(m.rcm_order())
    A = O.Block(*mo.arrays())
    f = B.Factor.ic2(mo, _fct_params(B, tau))
    ia, ja, a = f.arrays(1, A.nsup)
    ri, rj, ra = O.ref_ich2(A, tau1=tau[0], tau2=tau[1], memory=10.0)
    assert np.array_equal(ia, ri) and np.array_equal(ja, rj)
    assert np.array_equal(a, ra)


@pytest.mark.parametrize("bla,skew,tau", [(24, 0.3, (1e-2, 1e-4)), (6, 0.3, (1e-2, 1e-4)), (9, 0.0, (0.05, 5e-3))])
def test_ilu2_bitwise_equals_reference(B, O, bla, skew, tau):
    """The product's ILU2 factors L (by columns) and U (by rows) equal the
    reference's CombineLU + CSMatrixRS::Ilu2 (ilu2rs.cpp:1458-1817) with its
    LAPACK 3.0 DGESVD, bit for bit."""
    need_ref(O)
    m, _ = B.gen_seepage3d(4, 4, 3, bla, skew, 43)
    mo = m.reorder(m.rcm_order())
    A = O.Block(*mo.arrays())
    f = B.Factor.ilu2(mo, _fct_params(B, tau))
    (li, lj, la), (ui, uj, ua) = O.ref_ilu2(A, tau1=tau[0], tau2=tau[1], memory=10.0)
    for which, (ri, rj, ra) in ((0, (li, lj, la)), (1, (ui, uj, ua))):
        ia, ja, a = f.arrays(which, A.nsup)
        assert np.array_equal(ia, ri) and np.array_equal(ja, rj), which
        assert np.array_equal(a, ra), which


# ---- binary files in the reference's layout ----

@pytest.mark.parametrize("kind", ["heat-upper", "seepage-full"])
def test_matrix_binary_file_matches_reference(B, O, kind, tmp_path):
    """bcm_matrix_write_bin writes the same bytes as the reference's
    CSMatrixRS::WriteMtrBin (smatrixrs.cpp:3210-3227, compiled in
    oracle/_ref) for the same matrix, and bcm_matrix_read_bin reads the
    reference's file back to the same arrays."""
    need_ref(O)
    if kind == "heat-upper":
        m, _ = B.gen_heat2d(7, 6, 9, 3)
        symm = True
    else:
        m, _ = B.gen_seepage3d(4, 3, 3, 6, 0.3, 4)
        symm = False
    A = O.Block(*m.arrays())
    ours, theirs = str(tmp_path / "ours.bin"), str(tmp_path / "ref.bin")
    m.write_bin(ours)
    O.ref_write_mtr_bin(A, theirs)
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    back = B.Matrix.read_bin(theirs, symmetric=symm)
    for u, v in zip(back.arrays(), m.arrays()):
        assert np.array_equal(u, v)
    assert back.info()["symmetric"] == int(symm)


def test_vector_binary_file_matches_reference(B, O, tmp_path):
    """bcm_vector_write_bin writes CSVector::WriteVectBin's bytes
    (svector.cpp:537-546), the reference's ReadVectBin (svector.cpp:488-507)
    reads them, and bcm_vector_read_bin reads the reference's file."""
    need_ref(O)
    v = np.asfortranarray(np.random.default_rng(9).standard_normal((123, 3)))
    ours, theirs = str(tmp_path / "v_ours.bin"), str(tmp_path / "v_ref.bin")
    B.vector_write_bin(ours, v)
    O.ref_write_vect_bin(v, theirs)
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    assert np.array_equal(O.ref_read_vect_bin(ours), v)
    assert np.array_equal(B.vector_read_bin(theirs), v)


def test_binary_files_reject_truncation(B, tmp_path):
    m, _ = B.gen_seepage3d(3, 3, 2, 3, 0.3, 5)
    p = str(tmp_path / "m.bin")
    m.write_bin(p)
    data = open(p, "rb").read()
    open(p, "wb").write(data[: len(data) - 8])
    with pytest.raises(B.BcmError) as e:
        B.Matrix.read_bin(p)
    assert e.value.code == B.E_ARG
    with pytest.raises(B.BcmError):
        B.vector_read_bin(str(tmp_path / "missing.bin"))
