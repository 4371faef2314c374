"""Builders of the small systems the tests share (sizes the oracle finishes in seconds)."""

import numpy as np


def heat(B, O, nx=8, ny=7, bla=9, seed=3, **fct):
    """2D heat family (config-0 shape): RCM-ordered upper storage, IC2 factor."""
    m, rhs = B.gen_heat2d(nx, ny, bla, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    f = B.Factor.ic2(mo, B.default_params(**fct) if fct else None)
    A = O.Block(*mo.arrays())
    U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    b = mo.reorder_vector(order, rhs)
    return dict(m=m, mo=mo, f=f, A=A, U=U, L=U, b=b, order=order, rhs=rhs)


def seepage(B, O, nx=5, ny=4, nz=3, bla=6, skew=0.3, seed=11, **fct):
    """3D nonsymmetric family (config-1 shape): RCM-ordered full storage, ILU2 factors."""
    m, rhs = B.gen_seepage3d(nx, ny, nz, bla, skew, seed)
    order = m.rcm_order()
    mo = m.reorder(order)
    f = B.Factor.ilu2(mo, B.default_params(**fct) if fct else None)
    A = O.Block(*mo.arrays())
    L = O.Block(A.sprnd, *f.arrays(0, A.nsup))
    U = O.Block(A.sprnd, *f.arrays(1, A.nsup))
    b = mo.reorder_vector(order, rhs)
    return dict(m=m, mo=mo, f=f, A=A, L=L, U=U, b=b, order=order, rhs=rhs)


def dense_factor(blk):
    """Dense image of a factor with inverted diagonal blocks (diagonal re-inverted)."""
    n = blk.n
    d = np.zeros((n, n))
    off = 0
    sp = blk.sprnd
    for i in range(blk.nsup):
        bi = sp[i + 1] - sp[i]
        for j in range(blk.ia[i], blk.ia[i + 1]):
            c = blk.ja[j]
            bj = sp[c + 1] - sp[c]
            b = blk.a[off:off + bi * bj].reshape(bj, bi).T
            if c == i:
                b = np.linalg.inv(b)
            d[sp[i]:sp[i + 1], sp[c]:sp[c + 1]] = b
            off += bi * bj
    return d
