"""Local operators on one axis of a state tensor (src/localops.jl:11-188)."""
from __future__ import annotations

import numpy as np


class LocalOp:
    """An operator on one axis: DenseOp, DiagonalOp, BandedOp or ShiftOp."""


class DenseOp(LocalOp):
    """``DenseOp(M)``: a square dense matrix (src/localops.jl:18-26)."""

    def __init__(self, M):
        M = np.asarray(M)
        if M.ndim != 2 or M.shape[0] != M.shape[1]:
            from ._lib import DimensionMismatch
            raise DimensionMismatch(f"DenseOp requires a square matrix, got {M.shape}")
        self.M = M.astype(np.complex128 if np.iscomplexobj(M) else np.float64)

    @property
    def n(self):
        return self.M.shape[0]


class DiagonalOp(LocalOp):
    """``DiagonalOp(d)`` (src/localops.jl:33-35)."""

    def __init__(self, d):
        d = np.asarray(d)
        self.d = d.astype(np.complex128 if np.iscomplexobj(d) else np.float64)

    @property
    def n(self):
        return len(self.d)


class BandedOp(LocalOp):
    """``BandedOp(n, offsets, weights)``: Y[i] = sum_b w_b[i] X[i + o_b] over
    0 <= i + o_b < n (src/localops.jl:43-55)."""

    def __init__(self, n, offsets, weights):
        from ._lib import ArgumentError, DimensionMismatch
        if n < 0:
            raise ArgumentError(f"BandedOp requires n >= 0, got {n}")
        if len(offsets) != len(weights):
            raise DimensionMismatch(f"BandedOp: {len(offsets)} offsets but {len(weights)} weight vectors")
        ws = [np.asarray(w) for w in weights]
        if any(len(w) != n for w in ws):
            raise DimensionMismatch(f"BandedOp: every weights[b] must have length n = {n}")
        cplx = any(np.iscomplexobj(w) for w in ws)
        self.n = int(n)
        self.offsets = [int(o) for o in offsets]
        self.weights = [w.astype(np.complex128 if cplx else np.float64) for w in ws]


class ShiftOp(LocalOp):
    """``ShiftOp(n, s, factor)``: Y[i + s] = factor X[i] (src/localops.jl:63-71)."""

    def __init__(self, n, s, factor):
        from ._lib import ArgumentError
        if n < 0:
            raise ArgumentError(f"ShiftOp requires n >= 0, got {n}")
        self.n, self.s, self.factor = int(n), int(s), factor


def dimension(op):
    return op.n


def is_complex(op):
    if isinstance(op, DenseOp):
        return np.iscomplexobj(op.M)
    if isinstance(op, DiagonalOp):
        return np.iscomplexobj(op.d)
    if isinstance(op, BandedOp):
        return any(np.iscomplexobj(w) for w in op.weights)
    return isinstance(op.factor, complex) or np.iscomplexobj(op.factor)


def dense(op):
    """``dense(op)`` (src/localops.jl:115-130)."""
    if isinstance(op, DenseOp):
        return op.M.copy()
    if isinstance(op, DiagonalOp):
        return np.diag(op.d)
    n = op.n
    if isinstance(op, BandedOp):
        A = np.zeros((n, n), dtype=np.complex128 if is_complex(op) else np.float64)
        for o, w in zip(op.offsets, op.weights):
            i = np.arange(max(0, -o), min(n, n - o))
            A[i, i + o] += w[i]
        return A
    A = np.zeros((n, n), dtype=np.complex128 if is_complex(op) else np.float64)
    j = np.arange(max(0, -op.s), min(n, n - op.s))
    A[j + op.s, j] = op.factor
    return A


def realpart(op):
    """``realpart(op)`` (src/localops.jl:157-167)."""
    if isinstance(op, DenseOp):
        return DenseOp(np.real(op.M))
    if isinstance(op, DiagonalOp):
        return DiagonalOp(np.real(op.d))
    if isinstance(op, BandedOp):
        return BandedOp(op.n, list(op.offsets), [np.real(w) for w in op.weights])
    return ShiftOp(op.n, op.s, float(np.real(op.factor)))


def imagpart(op):
    """``imagpart(op)`` (src/localops.jl:157-167)."""
    if isinstance(op, DenseOp):
        return DenseOp(np.imag(op.M))
    if isinstance(op, DiagonalOp):
        return DiagonalOp(np.imag(op.d))
    if isinstance(op, BandedOp):
        return BandedOp(op.n, list(op.offsets), [np.imag(w) for w in op.weights])
    return ShiftOp(op.n, op.s, float(np.imag(op.factor)))


def same_op(a, b):
    """Structural equality of two local operators' data."""
    if type(a) is not type(b):
        return False
    if isinstance(a, DenseOp):
        return a.M.shape == b.M.shape and np.array_equal(a.M, b.M)
    if isinstance(a, DiagonalOp):
        return np.array_equal(a.d, b.d)
    if isinstance(a, BandedOp):
        return a.n == b.n and a.offsets == b.offsets and all(np.array_equal(x, y) for x, y in zip(a.weights, b.weights))
    return a.n == b.n and a.s == b.s and a.factor == b.factor
