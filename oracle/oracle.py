"""ORACLE - test infrastructure only.

ctypes access to the CPU oracle (liboracle.so, a plain-C restatement of the
reference's solve phase) and to the reference's own compiled solve-phase code
(_ref/libbslib_ref.so, built from the Bslib sources by Makefile.ref). Only the
tests, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg use this
module, as the checker and as the timed CPU baseline.

Block matrices are passed as ``(sprnd, ia, ja, a)`` numpy arrays with blocks
contiguous in ja order (column-major), as the C ABI of libbcmhip exports them.
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libbslib_ref.so")

_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
# Callback types of the reference's abstract solvers (smatrix.h:14-24): CNSUP
# and CSPRNDS fill ints, CIAJA returns two int arrays, the complex row /
# right-hand side / guess / solution callbacks pass interleaved re/im doubles.
CB_INTS = C.CFUNCTYPE(None, C.c_void_p, _IP)
CB_IAJA = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(_IP), C.POINTER(_IP))
CB_CROW = C.CFUNCTYPE(None, C.c_void_p, C.c_int, _DP)
_LP = C.POINTER(C.c_longlong)


class Bsr(C.Structure):
    _fields_ = [("nsup", C.c_int), ("sprnd", _IP), ("ia", _IP), ("ja", _IP), ("bsa", _LP), ("a", _DP)]


class Result(C.Structure):
    _fields_ = [("nitperf", C.c_int), ("updates", C.c_int), ("resfin", C.c_double), ("breakdown", C.c_int)]


_orc: Optional[C.CDLL] = None
_ref: Optional[C.CDLL] = None


def build() -> None:
    """Compiles the oracle (and the reference library when its sources exist)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE], check=True)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build()
        l = C.CDLL(ORACLE_SO)
        for name in ("orc_mvma", "orc_mvmat", "orc_solvel", "orc_solveu"):
            getattr(l, name).argtypes = [C.POINTER(Bsr), _DP, _DP, C.c_int]
        l.orc_mvmasymm.argtypes = [C.POINTER(Bsr), _DP, _DP]
        l.orc_scprod.argtypes = [C.c_int, _DP, _DP]
        l.orc_scprod.restype = C.c_double
        l.orc_cg.argtypes = [C.POINTER(Bsr), C.POINTER(Bsr), _DP, _DP, C.c_int, C.c_double, C.c_int, _DP]
        l.orc_cg.restype = Result
        l.orc_cg_par.argtypes = [C.POINTER(Bsr), C.POINTER(Bsr), _DP, _DP, C.c_int, C.c_double, C.c_int, C.c_int, _DP]
        l.orc_cg_par.restype = Result
        l.orc_lanczos.argtypes = [C.POINTER(Bsr)] * 3 + [_DP, _DP, C.c_int, C.c_double, C.c_int, _DP]
        l.orc_lanczos.restype = Result
        l.orc_bicgstab.argtypes = [C.POINTER(Bsr)] * 3 + [_DP, _DP, C.c_int, C.c_double, _DP]
        l.orc_bicgstab.restype = Result
        l.orc_fgmres.argtypes = [C.POINTER(Bsr)] * 3 + [_DP, _DP, C.c_int, C.c_int, C.c_double, C.c_int, _DP]
        l.orc_block_gmres.restype = Result
        l.orc_block_gmres.argtypes = [C.POINTER(Bsr)] * 3 + [C.c_int, C.c_int, _DP, _DP, C.c_int, C.c_double,
                                                            C.c_int, _DP, _DP, _DP]
        l.orc_fgmres.restype = Result
        l.orc_block_lanczos.argtypes = [C.c_int] + [C.POINTER(Bsr)] * 3 + [C.c_int, _DP, _DP, C.c_int, C.c_double,
                                                                           C.c_int, _DP]
        l.orc_block_lanczos.restype = Result
        _orc = l
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> C.CDLL:
    """The reference's compiled solve-phase code (lazy binding: MPI/LAPACK paths are never called)."""
    global _ref
    if _ref is None:
        l = C.CDLL(REF_SO, mode=os.RTLD_LAZY)
        mat = [C.c_int, _IP, _IP, _IP, _DP]
        for name in ("ref_mvma", "ref_mvmat", "ref_solvel", "ref_solveu"):
            getattr(l, name).argtypes = mat + [_DP, _DP, C.c_int]
        l.ref_mvmasymm.argtypes = mat + [_DP, _DP]
        l.ref_scprod.argtypes = [C.c_int, _DP, _DP]
        l.ref_scprod.restype = C.c_double
        l.ref_cg.argtypes = mat + [_IP, _IP, _DP, _DP, _DP, C.c_int, C.c_double, C.c_int, _DP, _IP, _DP]
        l.ref_lanczos.argtypes = mat + [_IP, _IP, _DP, _IP, _IP, _DP, _DP, _DP, C.c_int, C.c_double, C.c_int,
                                        _DP, _IP, _DP]
        l.ref_genrcm.argtypes = [C.c_int, _IP, _IP, _IP]
        l.ref_order_prf.argtypes = mat + [_IP]
        l.ref_find_separator.argtypes = [C.c_int, _IP, _IP, _IP, _IP]
        l.ref_cmvma.argtypes = mat + [_DP, _DP, C.c_int]
        l.ref_csolve.argtypes = [C.c_int] + mat + [_DP, _DP, C.c_int]
        l.ref_cilu2.argtypes = mat + [C.c_double] * 6 + [C.c_int, C.c_char_p, C.c_int]
        l.ref_zgesvd.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _DP, C.c_int, _DP, _IP]
        l.ref_alloc_ints.argtypes = [C.c_int]
        l.ref_cblock_gmres.argtypes = mat + [_DP, _DP, C.c_int, C.c_double]
        l.ref_cblock_gmres.restype = C.c_double
        l.ref_alloc_ints.restype = _IP
        l.ref_abstract_solver_c.argtypes = ([C.c_void_p, C.c_int, C.c_int] + [C.c_double] * 6 + [C.c_int, C.c_double]
                                            + [CB_INTS, CB_INTS, CB_IAJA, CB_CROW, CB_CROW, CB_CROW, CB_CROW]
                                            + [C.c_char_p, C.c_int])
        l.ref_cfct_take.argtypes = [C.c_int, _IP, _IP, _IP, _IP, _DP]
        l.ref_ordmtr.argtypes = [C.c_int] + mat + [_IP, _IP, _IP, _IP, _IP, _DP]
        l.ref_combinelu.argtypes = [C.c_int] + mat + [_IP, _IP, _IP, _IP, _DP]
        l.ref_ich2.argtypes = mat + [C.c_double] * 4 + [C.c_char_p, C.c_int]
        l.ref_ilu2.argtypes = mat + [C.c_double] * 6 + [C.c_char_p, C.c_int]
        l.ref_fct_take.argtypes = [C.c_int, _IP, _IP, _IP, _IP, _DP]
        l.ref_block_lanczos.argtypes = [C.c_int] + mat + [_IP, _IP, _DP, _IP, _IP, _DP, C.c_int, _DP, _DP,
                                                           C.c_int, C.c_double, C.c_int, _DP]
        l.ref_write_mtr_bin.argtypes = mat + [C.c_char_p]
        l.ref_write_vect_bin.argtypes = [C.c_int, C.c_int, _DP, C.c_char_p]
        l.ref_read_vect_bin.argtypes = [C.c_char_p, _IP, _IP, _DP]
        l.ref_dsyev.argtypes = [C.c_int, _DP, _DP, _DP, C.c_int, _IP]
        l.ref_dgesvd.argtypes = [C.c_int, _DP, _DP, _DP, _DP, _DP, C.c_int, _IP]
        _ref = l
    return _ref


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Block:
    """A block-CSR matrix as numpy arrays plus its C view for the oracle."""

    def __init__(self, sprnd, ia, ja, a):
        self.sprnd, self.ia, self.ja, self.a = _i(sprnd), _i(ia), _i(ja), _d(a)
        self.nsup = len(self.sprnd) - 1
        self.n = int(self.sprnd[-1])
        bla = np.diff(self.sprnd)
        rows = np.repeat(np.arange(self.nsup), np.diff(self.ia))
        sizes = bla[rows].astype(np.int64) * bla[self.ja].astype(np.int64)
        self.bsa = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(sizes) else np.zeros(0, np.int64)
        self.c = Bsr(self.nsup, self.sprnd.ctypes.data_as(_IP), self.ia.ctypes.data_as(_IP),
                     self.ja.ctypes.data_as(_IP), self.bsa.ctypes.data_as(_LP), self.a.ctypes.data_as(_DP))

    def args(self):
        return (self.nsup, self.sprnd.ctypes.data_as(_IP), self.ia.ctypes.data_as(_IP),
                self.ja.ctypes.data_as(_IP), self.a.ctypes.data_as(_DP))


def _dp(x):
    return x.ctypes.data_as(_DP)


def _vec(x, n, nrhs):
    return np.asfortranarray(np.asarray(x, dtype=np.float64).reshape((n, nrhs), order="F"))


# ---- oracle (restatement) ----

def mvma(m: Block, x, nrhs=1):
    x = _vec(x, m.n, nrhs); y = np.zeros_like(x)
    orc().orc_mvma(C.byref(m.c), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def mvmat(m: Block, x, nrhs=1):
    x = _vec(x, m.n, nrhs); y = np.zeros_like(x)
    orc().orc_mvmat(C.byref(m.c), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def mvmasymm(m: Block, x):
    x = _d(x); y = np.zeros_like(x)
    orc().orc_mvmasymm(C.byref(m.c), _dp(x), _dp(y))
    return y


def solvel(f: Block, x, nrhs=1):
    x = _vec(x, f.n, nrhs); y = np.zeros_like(x)
    orc().orc_solvel(C.byref(f.c), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def solveu(f: Block, x, nrhs=1):
    x = _vec(x, f.n, nrhs); y = np.zeros_like(x)
    orc().orc_solveu(C.byref(f.c), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def cg(a: Block, u: Block, b, x0=None, niter=1000, eps=1e-10, ittype=3):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    r = orc().orc_cg(C.byref(a.c), C.byref(u.c), _dp(x0), _dp(b), niter, eps, ittype, _dp(x))
    return x, r


def cg_par(a: Block, u: Block, b, x0=None, niter=1000, eps=1e-10, sttype=1, apply_last=True):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    r = orc().orc_cg_par(C.byref(a.c), C.byref(u.c), _dp(x0), _dp(b), niter, eps, sttype, int(apply_last), _dp(x))
    return x, r


def lanczos(a: Block, l: Block, u: Block, b, x0=None, niter=1000, eps=1e-10, sttype=1):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    r = orc().orc_lanczos(C.byref(a.c), C.byref(l.c), C.byref(u.c), _dp(x0), _dp(b), niter, eps, sttype, _dp(x))
    return x, r


def bicgstab(a: Block, l: Block, u: Block, b, x0=None, niter=1000, eps=1e-10):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    r = orc().orc_bicgstab(C.byref(a.c), C.byref(l.c), C.byref(u.c), _dp(x0), _dp(b), niter, eps, _dp(x))
    return x, r


def fgmres(a: Block, l: Block, u: Block, b, x0=None, m=30, niter=1000, eps=1e-10, symm=False):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    r = orc().orc_fgmres(C.byref(a.c), C.byref(l.c), C.byref(u.c), _dp(x0), _dp(b), m, niter, eps, int(symm),
                         _dp(x))
    return x, r


def block_lanczos(variant: int, a: Block, l: Block, u: Block, b, x0=None, niter=200, eps=1e-10, ichk=5):
    """Block Lanczos of the reference (lanczos.cpp:932-3225): variant 0 Right,
    1 Left, 2 Central preconditioning, on the n x nrhs block b. Returns (x, result)."""
    b = np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(a.n, -1, order="F"))
    nrhs = b.shape[1]
    x0 = np.zeros_like(b) if x0 is None else np.asfortranarray(np.asarray(x0, dtype=np.float64).reshape(a.n, -1,
                                                                                                       order="F"))
    x = np.zeros_like(b)
    r = orc().orc_block_lanczos(variant, C.byref(a.c), C.byref(l.c), C.byref(u.c), nrhs, _dp(x0), _dp(b), niter,
                                eps, ichk, _dp(x))
    return x, r


def block_gmres(a: Block, l: Block, u: Block, b, x0=None, niter=50, eps=1e-10, sttype=1, symm=False):
    """Block GMRES (gmres.cpp:337-1038) on the n x nrhs block b; returns
    (x, result, hist, fin): hist[k, c] the recursion's residual norm of column
    c after step k (zero past the exit), fin[c] = ||b_c - A x_c||."""
    b = np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(a.n, -1, order="F"))
    nrhs = b.shape[1]
    assert 1 <= nrhs <= 32
    x0 = np.zeros_like(b) if x0 is None else np.asfortranarray(np.asarray(x0, dtype=np.float64).reshape(b.shape, order="F"))
    x = np.zeros_like(b)
    hist = np.zeros((niter, nrhs))
    fin = np.zeros(nrhs)
    r = orc().orc_block_gmres(C.byref(a.c), C.byref(l.c), C.byref(u.c), int(symm), nrhs, _dp(x0), _dp(b), niter,
                              eps, sttype, _dp(x), _dp(hist), _dp(fin))
    return x, r, hist, fin


# ---- the reference's compiled code ----

def ref_mvma(m: Block, x, nrhs=1):
    x = _vec(x, m.n, nrhs); y = np.zeros_like(x)
    ref().ref_mvma(*m.args(), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def ref_mvmat(m: Block, x, nrhs=1):
    x = _vec(x, m.n, nrhs); y = np.zeros_like(x)
    ref().ref_mvmat(*m.args(), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def ref_mvmasymm(m: Block, x):
    x = _d(x); y = np.zeros_like(x)
    ref().ref_mvmasymm(*m.args(), _dp(x), _dp(y))
    return y


def ref_solvel(f: Block, x, nrhs=1):
    x = _vec(x, f.n, nrhs); y = np.zeros_like(x)
    ref().ref_solvel(*f.args(), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def ref_solveu(f: Block, x, nrhs=1):
    x = _vec(x, f.n, nrhs); y = np.zeros_like(x)
    ref().ref_solveu(*f.args(), _dp(x), _dp(y), nrhs)
    return y.reshape(-1, order="F") if nrhs == 1 else y


def ref_cg(a: Block, u: Block, b, x0=None, niter=1000, eps=1e-10, ittype=0) -> Tuple[np.ndarray, int, float]:
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    nit = C.c_int(); res = C.c_double()
    ref().ref_cg(*a.args(), u.ia.ctypes.data_as(_IP), u.ja.ctypes.data_as(_IP), u.a.ctypes.data_as(_DP),
                 _dp(x0), _dp(b), niter, eps, ittype, _dp(x), C.byref(nit), C.byref(res))
    return x, nit.value, res.value


def ref_lanczos(a: Block, l: Block, u: Block, b, x0=None, niter=1000, eps=1e-10, sttype=1):
    b = _d(b); x0 = np.zeros_like(b) if x0 is None else _d(x0); x = np.zeros_like(b)
    nit = C.c_int(); res = C.c_double()
    ref().ref_lanczos(*a.args(), l.ia.ctypes.data_as(_IP), l.ja.ctypes.data_as(_IP), l.a.ctypes.data_as(_DP),
                      u.ia.ctypes.data_as(_IP), u.ja.ctypes.data_as(_IP), u.a.ctypes.data_as(_DP),
                      _dp(x0), _dp(b), niter, eps, sttype, _dp(x), C.byref(nit), C.byref(res))
    return x, nit.value, res.value


def ref_genrcm(n, xadj0, adj0):
    """genrcm on a 0-based adjacency; returns the 0-based permutation perm[new] = old."""
    xadj = _i(np.asarray(xadj0) + 1); adj = _i(np.asarray(adj0) + 1); perm = np.zeros(n, np.int32)
    ref().ref_genrcm(n, xadj.ctypes.data_as(_IP), adj.ctypes.data_as(_IP), perm.ctypes.data_as(_IP))
    return perm - 1


def ref_order_prf(m: Block):
    order = np.zeros(m.nsup, np.int32)
    ref().ref_order_prf(*m.args(), order.ctypes.data_as(_IP))
    return order


def _cargs(sprnd, ia, ja, a):
    sprnd = np.ascontiguousarray(sprnd, np.int32)
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    av = np.ascontiguousarray(np.asarray(a, np.complex128)).view(np.float64)
    return (len(sprnd) - 1, sprnd.ctypes.data_as(_IP), ia.ctypes.data_as(_IP), ja.ctypes.data_as(_IP),
            av.ctypes.data_as(_DP)), (sprnd, ia, ja, av)


def _cvec(x, nrhs):
    return np.ascontiguousarray(np.asfortranarray(np.asarray(x, np.complex128).reshape(-1, nrhs)).ravel(order="F"))


def ref_cmvma(sprnd, ia, ja, a, x, nrhs=1):
    """The reference's complex CSMatrixCS::MvmA (mvmc.cpp:160): y = A x."""
    args, keep = _cargs(sprnd, ia, ja, a)
    xv = _cvec(x, nrhs)
    y = np.zeros_like(xv)
    ref().ref_cmvma(*args, xv.view(np.float64).ctypes.data_as(_DP), y.view(np.float64).ctypes.data_as(_DP), nrhs)
    return y.reshape(np.asarray(x).shape, order="F")


def ref_csolve(which, sprnd, ia, ja, a, x, nrhs=1):
    """The reference's complex SolveL (which 0, L by columns, mvmc.cpp:1723)
    or SolveU (which 1, U by rows, mvmc.cpp:2506)."""
    args, keep = _cargs(sprnd, ia, ja, a)
    xv = _cvec(x, nrhs)
    y = np.zeros_like(xv)
    ref().ref_csolve(which, *args, xv.view(np.float64).ctypes.data_as(_DP), y.view(np.float64).ctypes.data_as(_DP),
                     nrhs)
    return y.reshape(np.asarray(x).shape, order="F")


def ref_cilu2(sprnd, ia, ja, a, tau1=1e-2, tau2=1e-4, theta=0.1, sclmin=1e-8, pivmin=1e-8, memory=10.0, fcttyp=1):
    """The reference's CombineLU + complex CSMatrixCS::Ilu2 (ilu2cs.cpp):
    ((lia, lja, la), (uia, uja, ua)) with complex values, L by columns."""
    args, keep = _cargs(sprnd, ia, ja, a)
    msg = C.create_string_buffer(256)
    if ref().ref_cilu2(*args, tau1, tau2, theta, sclmin, pivmin, memory, fcttyp, msg, 256):
        raise RefFactorError(msg.value.decode(errors="replace"))
    nsup = len(keep[0]) - 1
    out = []
    for which in (0, 1):
        nzja, nza = C.c_int(), C.c_int()
        ref().ref_cfct_take(which, C.byref(nzja), C.byref(nza), None, None, None)
        fia = np.zeros(nsup + 1, np.int32)
        fja = np.zeros(nzja.value, np.int32)
        fa = np.zeros(nza.value, np.complex128)
        ref().ref_cfct_take(which, C.byref(nzja), C.byref(nza), fia.ctypes.data_as(_IP), fja.ctypes.data_as(_IP),
                            fa.view(np.float64).ctypes.data_as(_DP))
        out.append((fia, fja, fa))
    return tuple(out)


def ref_zgesvd(a):
    """zgesvd_("A", "A") through lapackd.cpp -> lapackdm.f -> ZGESVD_SM: (s, U, VT), complex column-major."""
    n = a.shape[0]
    A = np.array(a, dtype=np.complex128, order="F", copy=True)
    s = np.zeros(n)
    U = np.zeros((n, n), np.complex128, order="F")
    VT = np.zeros((n, n), np.complex128, order="F")
    lw = 10 * n
    work = np.zeros(lw, np.complex128)
    rwork = np.zeros(lw)
    info = C.c_int()
    ref().ref_zgesvd(n, A.ctypes.data_as(_DP), s.ctypes.data_as(_DP), U.ctypes.data_as(_DP),
                     VT.ctypes.data_as(_DP), work.ctypes.data_as(_DP), lw, rwork.ctypes.data_as(_DP), C.byref(info))
    return s, U, VT, info.value


def ref_cblock_gmres(sprnd, ia, ja, a, b, x0=None, niter=100, eps=0.0):
    """The reference's complex CSVectorC::BlockGMRES (gmres.cpp:1044) of one
    right-hand side with the factors of the last ref_cilu2 call: (x, host
    seconds of the call)."""
    args, keep = _cargs(sprnd, ia, ja, a)
    bv = np.ascontiguousarray(b, np.complex128)
    x = np.zeros_like(bv) if x0 is None else np.array(x0, np.complex128)
    sec = ref().ref_cblock_gmres(*args, bv.ctypes.data_as(_DP), x.ctypes.data_as(_DP), niter, eps)
    return x, sec


def ref_alloc_ints(n):
    """int[n] from the reference library's operator new[], for a CIAJA callback."""
    return ref().ref_alloc_ints(n)


def ref_abstract_solver_c(fnsup, fsprnds, fiaja, fblockrow, frhs, finiguess, fstoresol, ordtype=1, fcttyp=1,
                          tau1=1e-2, tau2=1e-4, theta=0.1, sclmin=1e-8, pivmin=1e-8, memory=10.0, niter=1000,
                          eps=1e-9):
    """The reference's complex AbstractSolver (solvers.cpp:7103) on Python
    callbacks with the arguments of CNSUP, CSPRNDS, CIAJA (arrays from
    ref_alloc_ints), CBLOCKROWC, CRHSC, CINIGUESSC, CSTORESOLC (complex data
    as interleaved doubles)."""
    cbs = [CB_INTS(fnsup), CB_INTS(fsprnds), CB_IAJA(fiaja), CB_CROW(fblockrow), CB_CROW(frhs), CB_CROW(finiguess),
           CB_CROW(fstoresol)]
    msg = C.create_string_buffer(256)
    if ref().ref_abstract_solver_c(None, ordtype, fcttyp, tau1, tau2, theta, sclmin, pivmin, memory, niter, eps,
                                   *cbs, msg, 256):
        raise RefFactorError(msg.value.decode(errors="replace"))


def ref_find_separator(ia, ja):
    """The reference's CSMatrix::FindSeparator (smatrix.cpp:4355) on a pattern."""
    ia = np.ascontiguousarray(ia, np.int32)
    ja = np.ascontiguousarray(ja, np.int32)
    i1, i2 = C.c_int(), C.c_int()
    ref().ref_find_separator(len(ia) - 1, ia.ctypes.data_as(_IP), ja.ctypes.data_as(_IP), C.byref(i1), C.byref(i2))
    return i1.value, i2.value


def _take(fn, *args, nsup):
    nzja = C.c_int(); nza = C.c_int()
    fn(*args, C.byref(nzja), C.byref(nza), None, None, None)
    ia = np.zeros(nsup + 1, np.int32); ja = np.zeros(nzja.value, np.int32); a = np.zeros(nza.value)
    fn(*args, C.byref(nzja), C.byref(nza), ia.ctypes.data_as(_IP), ja.ctypes.data_as(_IP), a.ctypes.data_as(_DP))
    return ia, ja, a


def ref_ordmtr(m: Block, order, symmetric: bool):
    order = _i(order)
    return _take(ref().ref_ordmtr, int(symmetric), *m.args(), order.ctypes.data_as(_IP), nsup=m.nsup)


def ref_combinelu(m: Block, which: int):
    return _take(ref().ref_combinelu, which, *m.args(), nsup=m.nsup)


class RefFactorError(RuntimeError):
    pass


def ref_ich2(m: Block, tau1=1e-2, tau2=1e-4, theta=0.1, memory=10.0):
    """The reference's CSMatrixRS::Ich2 (ich2rs.cpp:1075) on an upper-stored
    matrix: the U factor (ia, ja, a) in the reference layout."""
    msg = C.create_string_buffer(256)
    if ref().ref_ich2(*m.args(), tau1, tau2, theta, memory, msg, 256):
        raise RefFactorError(msg.value.decode(errors="replace"))
    return _take(ref().ref_fct_take, 1, nsup=m.nsup)


def ref_ilu2(m: Block, tau1=1e-2, tau2=1e-4, theta=0.1, sclmin=1e-8, pivmin=1e-8, memory=10.0):
    """CombineLU + the reference's CSMatrixRS::Ilu2 (ilu2rs.cpp:1458) on a
    full-storage matrix: ((ia, ja, a) of L by columns, (ia, ja, a) of U by rows)."""
    msg = C.create_string_buffer(256)
    if ref().ref_ilu2(*m.args(), tau1, tau2, theta, sclmin, pivmin, memory, msg, 256):
        raise RefFactorError(msg.value.decode(errors="replace"))
    return _take(ref().ref_fct_take, 0, nsup=m.nsup), _take(ref().ref_fct_take, 1, nsup=m.nsup)


def ref_block_lanczos(variant: int, a: Block, l: Block, u: Block, b, x0=None, niter=200, eps=1e-10, ichk=5):
    """The reference's CSVector::BlockLanczosRight (0) / Left (1) / Central (2)
    (lanczos.cpp:932, 1719, 2467) on the n x nrhs block b."""
    b = np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(a.n, -1, order="F"))
    nrhs = b.shape[1]
    x0 = np.zeros_like(b) if x0 is None else np.asfortranarray(np.asarray(x0, dtype=np.float64).reshape(a.n, -1,
                                                                                                       order="F"))
    sol = np.zeros_like(b)
    ref().ref_block_lanczos(variant, *a.args(), l.ia.ctypes.data_as(_IP), l.ja.ctypes.data_as(_IP),
                            l.a.ctypes.data_as(_DP), u.ia.ctypes.data_as(_IP), u.ja.ctypes.data_as(_IP),
                            u.a.ctypes.data_as(_DP), nrhs, x0.ctypes.data_as(_DP), b.ctypes.data_as(_DP), niter, eps,
                            ichk, sol.ctypes.data_as(_DP))
    return sol


def ref_write_mtr_bin(m: Block, path: str):
    """CSMatrixRS::WriteMtrBin (smatrixrs.cpp:3210) of m to path."""
    ref().ref_write_mtr_bin(*m.args(), path.encode())


def ref_write_vect_bin(v, path: str):
    """CSVector::WriteVectBin (svector.cpp:537) of an n x nrhs array."""
    v = np.asfortranarray(np.asarray(v, dtype=np.float64).reshape(v.shape[0], -1, order="F"))
    ref().ref_write_vect_bin(v.shape[0], v.shape[1], v.ctypes.data_as(_DP), path.encode())


def ref_read_vect_bin(path: str):
    """CSVector::ReadVectBin (svector.cpp:488): n x nrhs array."""
    n, nrhs = C.c_int(), C.c_int()
    ref().ref_read_vect_bin(path.encode(), C.byref(n), C.byref(nrhs), None)
    v = np.zeros((n.value, nrhs.value), order="F")
    ref().ref_read_vect_bin(path.encode(), C.byref(n), C.byref(nrhs), v.ctypes.data_as(_DP))
    return v


def ref_dsyev(a):
    """dsyev_("V", "U") through lapackd.cpp -> lapackdm.f -> DSYEV_SM: (w, Z column-major)."""
    n = a.shape[0]
    A = np.array(a, dtype=np.float64, order="F", copy=True)
    w = np.zeros(n)
    lw = 10 * n
    work = np.zeros(lw)
    info = C.c_int()
    ref().ref_dsyev(n, A.ctypes.data_as(_DP), w.ctypes.data_as(_DP),
                    work.ctypes.data_as(_DP), lw, C.byref(info))
    return w, A, info.value


def ref_dgesvd(a):
    """dgesvd_("A", "A") through lapackd.cpp -> lapackdm.f -> DGESVD_SM: (s, U, VT), column-major."""
    n = a.shape[0]
    A = np.array(a, dtype=np.float64, order="F", copy=True)
    s = np.zeros(n)
    U = np.zeros((n, n), order="F")
    VT = np.zeros((n, n), order="F")
    lw = 10 * n
    work = np.zeros(lw)
    info = C.c_int()
    ref().ref_dgesvd(n, A.ctypes.data_as(_DP), s.ctypes.data_as(_DP), U.ctypes.data_as(_DP),
                     VT.ctypes.data_as(_DP), work.ctypes.data_as(_DP), lw, C.byref(info))
    return s, U, VT, info.value
