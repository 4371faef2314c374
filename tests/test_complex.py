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
