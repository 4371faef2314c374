"""Complex solve phase as real equivalents (CSMatrixCS, mvmc.cpp; bcm_cmatrix_create,
bcm_cfactor_create), on the CPU.

Pins: the product's real-equivalent matrix, applied with plain real
arithmetic, equals the reference's own complex CSMatrixCS::MvmA (mvmc.cpp:160,
compiled in oracle/_ref) to rounding; the reference's own complex ILU2
factors (CombineLU + CSMatrixCS::Ilu2, ilu2cs.cpp), converted by the product
and swept by the oracle's real SolveL / SolveU, equal the reference's complex
SolveL / SolveU (mvmc.cpp:1723, 2506) to rounding — which fixes the sign
pattern of the transposed L blocks and the plain U blocks. The complex
arithmetic accumulates (re, im) pairs where the real form sums separate
column runs, hence rounding-level, not bitwise, agreement.
"""

import numpy as np
import pytest

from test_host import need_ref


def complex_system(B, nx=4, ny=4, nz=3, bla=6, seed=5):
    """A nonsymmetric complex block system on a seepage pattern: real part one
    seepage matrix, imaginary part 0.4 x another with the same pattern."""
    m, _ = B.gen_seepage3d(nx, ny, nz, bla, 0.3, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    sp, ia, ja, a = mo.arrays()
    m2, _ = B.gen_seepage3d(nx, ny, nz, bla, 0.5, seed + 4)
    a2 = m2.reorder(order).arrays()[3]
    return sp, ia, ja, a + 0.4j * a2


def dense_complex(sp, ia, ja, ac):
    n = sp[-1]
    d = np.zeros((n, n), complex)
    off = 0
    for i in range(len(sp) - 1):
        for j in range(ia[i], ia[i + 1]):
            c = ja[j]
            bi, bj = sp[i + 1] - sp[i], sp[c + 1] - sp[c]
            d[sp[i]:sp[i + 1], sp[c]:sp[c + 1]] = ac[off:off + bi * bj].reshape(bj, bi).T
            off += bi * bj
    return d


def test_vector_layout_round_trip(B):
    sp = np.array([0, 3, 5, 9])
    rng = np.random.default_rng(1)
    z = rng.standard_normal((9, 2)) + 1j * rng.standard_normal((9, 2))
    x = B.complex_to_real(sp, z)
    assert x.shape == (18, 2)
    # block row 1 (sizes 2): [Re z3, Re z4, Im z3, Im z4]
    assert np.array_equal(x[6:10, 0], np.array([z[3, 0].real, z[4, 0].real, z[3, 0].imag, z[4, 0].imag]))
    assert np.array_equal(B.real_to_complex(sp, x), z)


def test_real_equivalent_matrix_is_the_complex_operator(B, O):
    sp, ia, ja, ac = complex_system(B)
    mc = B.Matrix.from_complex(sp, ia, ja, ac)
    inf = mc.info()
    assert inf["n"] == 2 * sp[-1] and not inf["symmetric"]
    rng = np.random.default_rng(3)
    z = rng.standard_normal(sp[-1]) + 1j * rng.standard_normal(sp[-1])
    y = B.real_to_complex(sp, mc.dense() @ B.complex_to_real(sp, z))
    assert np.abs(y - dense_complex(sp, ia, ja, ac) @ z).max() <= 1e-13 * np.abs(y).max()
    need_ref(O)
    yr = O.ref_cmvma(sp, ia, ja, ac, z)
    assert np.abs(y - yr).max() <= 1e-13 * np.abs(yr).max()


def test_factor_sweeps_equal_reference_complex_sweeps(B, O):
    need_ref(O)
    sp, ia, ja, ac = complex_system(B)
    L, U = O.ref_cilu2(sp, ia, ja, ac)
    f = B.Factor.from_complex(sp, L, U)
    mc = B.Matrix.from_complex(sp, ia, ja, ac)
    spr = mc.arrays()[0]
    nsup = len(sp) - 1
    rng = np.random.default_rng(4)
    for nrhs in (1, 3):
        z = rng.standard_normal((sp[-1], nrhs)) + 1j * rng.standard_normal((sp[-1], nrhs))
        xr = B.complex_to_real(sp, z)
        for which, fn, part in ((0, O.solvel, L), (1, O.solveu, U)):
            blk = O.Block(spr, *f.arrays(which, nsup))
            yr = np.asarray(fn(blk, np.asfortranarray(xr), nrhs)).reshape(xr.shape, order="F")
            yc = B.real_to_complex(sp, yr)
            yref = O.ref_csolve(which, sp, *part, z, nrhs)
            assert np.abs(yc - yref).max() <= 1e-13 * np.abs(yref).max(), (which, nrhs)


def test_complex_input_validation(B):
    sp = np.array([0, 2, 4])
    ia = np.array([0, 1, 2])
    ja = np.array([0, 5])
    with pytest.raises(B.BcmError):
        B.Matrix.from_complex(sp, ia, ja, np.ones(8, complex))
    # a factor row must start with its diagonal block
    ja_ok = np.array([1, 1])
    with pytest.raises(B.BcmError):
        B.Factor.from_complex(sp, (ia, ja_ok, np.ones(8, complex)), (ia, np.array([0, 1]), np.ones(8, complex)))


def _complex_dense_cases(n, rng):
    """The real SVD test's stress cases in complex arithmetic: random,
    repeated singular values in a random unitary basis, diagonal with ties,
    rank deficient, graded, scaled into ZLASCL's ranges, zero, identity, and
    a real matrix (imaginary parts exactly zero)."""
    g = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    rep = np.repeat(np.arange(1, (n + 2) // 3 + 1, dtype=float), 3)[:n]
    k = max(1, n // 2)
    return {
        "random": g,
        "repeated": (q * rep) @ q.conj().T,
        "diag-ties": np.diag(np.repeat([3.0 - 1j, 1.0 + 2j], (n + 1) // 2)[:n]),
        "rank-def": g[:, :k] @ g[:k, :],
        "graded": g * np.logspace(0, -12, n)[None, :],
        "tiny": g * 1e-160,
        "huge": g * 1e160,
        "zero": np.zeros((n, n), complex),
        "identity": np.eye(n, dtype=complex),
        "real": g.real.astype(complex),
    }


@pytest.mark.parametrize("n", [2, 3, 9, 12, 24, 32])
def test_zgesvd_matches_reference_lapack(B, O, n):
    """bcm_zgesvd (LAPACK 3.0 ZGESVD restated) against zgesvd_("A", "A") of
    the reference's own LAPACK build: singular values and both singular
    vector sets bit for bit (== also equates zeros of either sign)."""
    need_ref(O)
    rng = np.random.default_rng(300 + n)
    for name, a in _complex_dense_cases(n, rng).items():
        s, u, vt = B.zgesvd(a)
        sr, ur, vtr, info = O.ref_zgesvd(a)
        assert info == 0, name
        assert np.array_equal(s, sr) and np.array_equal(u, ur) and np.array_equal(vt, vtr), (n, name)
        if name == "random":  # and it is a singular value decomposition
            assert np.allclose((u * s) @ vt, a, atol=1e-12 * np.abs(a).max())
            assert np.allclose(u.conj().T @ u, np.eye(n), atol=1e-13)


def _blocked(dense, sp):
    """Block CSR (complex values, blocks column-major in ja order) of a dense
    complex matrix over the partition sp: every block with a nonzero entry."""
    nsup = len(sp) - 1
    ia, ja, vals = [0], [], []
    for i in range(nsup):
        for j in range(nsup):
            blk = dense[sp[i]:sp[i + 1], sp[j]:sp[j + 1]]
            if i == j or np.any(blk != 0):
                ja.append(j)
                vals.append(blk.T.ravel())  # column-major
        ia.append(len(ja))
    return np.array(sp), np.array(ia), np.array(ja), np.concatenate(vals)


def _variable_system(seed=8, n=60):
    """A nonsymmetric complex matrix of 60 unknowns with a banded plus
    scattered pattern, on the partition 3, 1, 4, 2, 5, ... (variable blocks,
    1 x 1 blocks included)."""
    rng = np.random.default_rng(seed)
    d = np.zeros((n, n), complex)
    for k in range(-4, 5):
        idx = np.arange(max(0, -k), min(n, n - k))
        d[idx, idx + k] = rng.standard_normal(len(idx)) + 0.5j * rng.standard_normal(len(idx))
    far = rng.integers(0, n, (40, 2))
    d[far[:, 0], far[:, 1]] += 0.3 * (rng.standard_normal(40) + 1j * rng.standard_normal(40))
    d += np.diag(6.0 + 2.0j + rng.standard_normal(n))
    sizes, tot = [], 0
    for s in [3, 1, 4, 2, 5, 1, 6, 3] * 10:
        if tot + s > n:
            break
        sizes.append(s)
        tot += s
    sizes.append(n - tot) if tot < n else None
    sp = np.concatenate([[0], np.cumsum(sizes)])
    return _blocked(d, sp)


@pytest.mark.parametrize("case", ["b6-tau", "b12-tau-loose", "b6-pattern", "b1-tau", "variable-tau",
                                  "variable-pattern", "b12-theta0"])
def test_complex_ilu2_bitwise_equals_reference(B, O, case):
    """bcm_cfactor_ilu2 (CombineLU + the restated complex ILU2 over the
    restated ZGESVD) against the reference's own CSMatrixCS::CombineLU + Ilu2
    (ilu2cs.cpp:1213): the real-equivalent images of both factors equal bit
    for bit, so the complex factors do (the image only copies and negates)."""
    need_ref(O)
    if case.startswith("variable"):
        sp, ia, ja, ac = _variable_system()
    else:
        bla = int(case.split("-")[0][1:])
        sp, ia, ja, ac = complex_system(B, 5, 4, 4, bla, seed=40 + bla)
    tau = (0.1, 0.01) if "loose" in case else (1e-2, 1e-4)
    fcttyp = 0 if "pattern" in case else 1
    theta = 0.0 if "theta0" in case else 0.1
    p = B.default_params(tau1=tau[0], tau2=tau[1], theta=theta, memory=10.0, fcttyp=fcttyp)
    f = B.Factor.ilu2_complex(sp, ia, ja, ac, p)
    L, U = O.ref_cilu2(sp, ia, ja, ac, tau1=tau[0], tau2=tau[1], theta=theta, memory=10.0, fcttyp=fcttyp)
    fr = B.Factor.from_complex(sp, L, U)
    nsup = len(sp) - 1
    for which in (0, 1):
        got, want = f.arrays(which, nsup), fr.arrays(which, nsup)
        assert all(np.array_equal(g, w) for g, w in zip(got, want)), (case, which)
    # density: stored values of L per value of CombineLU's L part, the upper half of the symmetrized
    # pattern (ilu2cs.cpp:1544-1546); a real-equivalent block holds 4 values per complex entry
    pat = {(i, j) for i in range(nsup) for j in ja[ia[i]:ia[i + 1]]}
    pat |= {(j, i) for i, j in pat}
    nza_l = sum((sp[i + 1] - sp[i]) * (sp[j + 1] - sp[j]) for i, j in pat if j >= i)
    assert p.density == len(f.arrays(0, nsup)[2]) / 4 / nza_l and p.timefct > 0
    if fcttyp == 0:  # the pattern variant keeps exactly the blocks of CombineLU's parts: U = upper half
        rows = [sorted(j for (r, j) in pat if r == i and j >= i) for i in range(nsup)]
        assert np.array_equal(U[0], np.cumsum([0] + [len(r) for r in rows]))
        assert np.array_equal(U[1], np.concatenate(rows))
