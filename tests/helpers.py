"""Shared helpers: build the same operator in the product (kfc) and in the oracle."""
import numpy as np

from oracle import kfc_oracle as O


def random_ops(rng, dims, nterms, cplx, kinds=("dense", "diag", "banded", "shift"), maxoff=None):
    """A random term list as (coeff, [(axis, kind, data)])."""
    terms = []
    for t in range(nterms):
        k = int(rng.integers(1, len(dims) + 1))
        axes = sorted(int(a) for a in rng.choice(len(dims), size=k, replace=False))
        fs = []
        for a in axes:
            d = dims[a]
            kind = kinds[(t + a) % len(kinds)]
            r = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if cplx else \
                (lambda *s: rng.standard_normal(s))
            if kind == "dense":
                fs.append((a, "dense", r(d, d)))
            elif kind == "diag":
                fs.append((a, "diag", r(d)))
            elif kind == "banded":
                offs = [-1, 0, 1] if maxoff == 1 else [-2, 0, 1, d - 1]
                fs.append((a, "banded", (offs, [r(d) for _ in offs])))
            else:
                s = int(rng.integers(-1, 2)) if maxoff == 1 else int(rng.integers(-d + 1, d))
                fs.append((a, "shift", (s, complex(*rng.standard_normal(2)) if cplx else float(rng.standard_normal()))))
        c = complex(*rng.standard_normal(2)) if cplx else float(rng.standard_normal())
        terms.append((c, fs))
    return terms


def oracle_terms(dims, terms):
    """The oracle's AxisTerms of a random_ops term list."""
    ts = []
    for c, fs in terms:
        ops = []
        for a, kind, data in fs:
            if kind == "dense":
                ops.append((a, O.DenseOp(data)))
            elif kind == "diag":
                ops.append((a, O.DiagonalOp(data)))
            elif kind == "banded":
                ops.append((a, O.BandedOp(dims[a], data[0], data[1])))
            else:
                ops.append((a, O.ShiftOp(dims[a], data[0], data[1])))
        ts.append(O.AxisTerm(c, ops))
    return ts


def to_oracle(dims, terms, constant=0.0):
    return O.kron_operator(dims, oracle_terms(dims, terms), constant)


def to_kfc(dims, terms, constant=0.0, block_capacity=1):
    import kfc
    ts = []
    for c, fs in terms:
        ops = []
        for a, kind, data in fs:
            if kind == "dense":
                ops.append((a, kfc.DenseOp(data)))
            elif kind == "diag":
                ops.append((a, kfc.DiagonalOp(data)))
            elif kind == "banded":
                ops.append((a, kfc.BandedOp(dims[a], data[0], data[1])))
            else:
                ops.append((a, kfc.ShiftOp(dims[a], data[0], data[1])))
        ts.append(kfc.AxisTerm(c, ops))
    return kfc.KronFreeOperator(dims, ts, constant=constant, block_capacity=block_capacity)


def oracle_from_kfc_terms(H):
    """The oracle operator of a kfc operator's own term list (same numbers)."""
    import kfc
    ts = []
    for t in H.terms:
        ops = []
        for a, op in t.factors:
            if isinstance(op, kfc.DenseOp):
                ops.append((a, O.DenseOp(op.M)))
            elif isinstance(op, kfc.DiagonalOp):
                ops.append((a, O.DiagonalOp(op.d)))
            elif isinstance(op, kfc.BandedOp):
                ops.append((a, O.BandedOp(op.n, op.offsets, op.weights)))
            else:
                ops.append((a, O.ShiftOp(op.n, op.s, op.factor)))
        ts.append(O.AxisTerm(t.coeff, ops))
    return O.kron_operator(H.dims, ts, H.constant)
