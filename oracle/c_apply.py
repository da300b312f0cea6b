"""ctypes front end of the oracle's C restatement (apply_chain.c).  TEST
INFRASTRUCTURE ONLY: tests/ and bench.py's cpu_baseline leg."""
import ctypes as C
import os

import numpy as np

from . import kfc_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "liboracle_apply.so")


class _F(C.Structure):
    _fields_ = [("axis", C.c_int32), ("kind", C.c_int32), ("d", C.c_int64), ("nb", C.c_int32),
                ("offs", C.POINTER(C.c_int64)), ("w", C.POINTER(C.c_double)), ("s", C.c_int64), ("f", C.c_double),
                ("M", C.POINTER(C.c_double)), ("dg", C.POINTER(C.c_double))]


class _T(C.Structure):
    _fields_ = [("c", C.c_double), ("nf", C.c_int32), ("f", C.POINTER(_F))]


_lib = None


def available():
    return os.path.exists(LIB)


def _load():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB)
        _lib.oracle_apply.restype = None
    return _lib


class CApply:
    """The reference's per-term chain apply of a real oracle KronOp, in C."""

    def __init__(self, H):
        if H.dtype != np.float64:
            raise ValueError("the C restatement serves real operators")
        self.H = H
        self.keep = []
        terms = (_T * max(1, len(H.terms)))()
        D = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        for ti, t in enumerate(H.terms):
            fs = (_F * len(t.factors))()
            for k, (a, op) in enumerate(t.factors):
                f = fs[k]
                f.axis, f.d = a, op.n
                if isinstance(op, O.DiagonalOp):
                    arr = np.ascontiguousarray(op.d, dtype=np.float64)
                    f.kind, f.dg = 0, D(arr)
                elif isinstance(op, O.BandedOp):
                    arr = np.ascontiguousarray(np.stack([np.real(w) for w in op.weights], 1).reshape(-1, order="F"))
                    offs = np.asarray(op.offsets, dtype=np.int64)
                    f.kind, f.nb, f.w, f.offs = 1, len(op.offsets), D(arr), offs.ctypes.data_as(C.POINTER(C.c_int64))
                    self.keep.append(offs)
                elif isinstance(op, O.ShiftOp):
                    arr = None
                    f.kind, f.s, f.f = 2, op.s, float(np.real(op.factor))
                else:
                    arr = np.ascontiguousarray(np.real(op.M).reshape(-1, order="F"))
                    f.kind, f.M = 3, D(arr)
                if arr is not None:
                    self.keep.append(arr)
            self.keep.append(fs)
            terms[ti].c, terms[ti].nf, terms[ti].f = float(np.real(t.coeff)), len(t.factors), fs
        self.terms = terms
        self.diag = np.ascontiguousarray(np.real(H.diag), dtype=np.float64)
        self.dims = np.asarray(H.dims, dtype=np.int64)
        self.b1 = np.empty(H.N)
        self.b2 = np.empty(H.N)

    def __call__(self, x, y=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x) if y is None else y
        D = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        _load().oracle_apply(C.c_int32(len(self.dims)), self.dims.ctypes.data_as(C.POINTER(C.c_int64)), D(self.diag),
                             C.c_int32(len(self.H.terms)), self.terms, D(x), D(y), D(self.b1), D(self.b2))
        return y
