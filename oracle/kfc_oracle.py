"""CPU oracle for the Kronecker-free apply and its eigensolver loops.

TEST INFRASTRUCTURE ONLY.  This module restates, in numpy, the algorithm of
KronFreeCircuits.jl for the hot path (local operators, basis port operators,
KronFreeOperator assembly and the per-term two-buffer apply chain, derived
quantities, circuit-term expansion and the real-node expansion, a small
tree materialization with Dense solves, and the linear solve on the complement
of a node's known states).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, and only as the checker: no product code path calls it.

Pinning (what makes each part trustworthy; the checks are in tests/test_oracle.py):
  * fd weights: exact rationals and moment identities (test/test_bases.jl:6, 47-54)
  * basis operators: closed forms (Fock ladders, charge shifts, Gaussian
    derivative accuracy, plane-wave symbols of the cyclic stencils)
  * apply chain: brute-force dense Kronecker sums (kron_reference,
    test/helpers.jl:10-24) on small dims
  * end to end: scqubits 4.3.1 golden energies in
    test/fixtures/scqubits_reference.json (transmon, fluxonium, zero-pi,
    full zero-pi, two transmons) at the reference's own tolerances
  * real expansion: real operator == complex operator on every config
  * solvers: numpy.linalg.eigh (LAPACK) on dense matrices
  * complement solve: the spectral solution on dense matrices (eigh)

Reference citations are ``src/...`` / ``test/...`` paths of KronFreeCircuits.jl.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field
from fractions import Fraction as Fr

import numpy as np

# ----------------------------------------------------------------------------------------
# Local operators (src/localops.jl:11-188)


@dataclass
class DenseOp:
    """``DenseOp(M)``: square dense matrix (src/localops.jl:18-26)."""
    M: np.ndarray

    @property
    def n(self):
        return self.M.shape[0]


@dataclass
class DiagonalOp:
    """``DiagonalOp(d)`` (src/localops.jl:33-35)."""
    d: np.ndarray

    @property
    def n(self):
        return len(self.d)


@dataclass
class BandedOp:
    """``BandedOp(n, offsets, weights)``: Y[i] = sum_b w_b[i] X[i+o_b] over valid
    i+o_b (src/localops.jl:43-55)."""
    n: int
    offsets: list
    weights: list


@dataclass
class ShiftOp:
    """``ShiftOp(n, s, f)``: Y[i+s] = f X[i] (src/localops.jl:63-71)."""
    n: int
    s: int
    factor: complex


def dense(op):
    """``dense(op)`` (src/localops.jl:115-130)."""
    if isinstance(op, DenseOp):
        return np.array(op.M)
    if isinstance(op, DiagonalOp):
        return np.diag(op.d)
    if isinstance(op, BandedOp):
        dt = np.result_type(*[np.asarray(w) for w in op.weights]) if op.weights else float
        A = np.zeros((op.n, op.n), dtype=dt)
        for o, w in zip(op.offsets, op.weights):
            for i in range(op.n):
                if 0 <= i + o < op.n:
                    A[i, i + o] += w[i]
        return A
    if isinstance(op, ShiftOp):
        A = np.zeros((op.n, op.n), dtype=np.result_type(op.factor, float))
        for i in range(op.n):
            if 0 <= i + op.s < op.n:
                A[i + op.s, i] = op.factor
        return A
    raise TypeError(op)


def diagonal(op):
    """``diagonal(op)`` (src/localops.jl:102-108)."""
    if isinstance(op, DenseOp):
        return np.diag(op.M).copy()
    if isinstance(op, DiagonalOp):
        return np.array(op.d)
    if isinstance(op, BandedOp):
        for o, w in zip(op.offsets, op.weights):
            if o == 0:
                return np.array(w)
        return np.zeros(op.n)
    if isinstance(op, ShiftOp):
        return np.full(op.n, op.factor) if op.s == 0 else np.zeros(op.n)
    raise TypeError(op)


def realpart(op):
    """``realpart`` (src/localops.jl:157-167)."""
    if isinstance(op, DenseOp):
        return DenseOp(np.real(op.M).astype(float))
    if isinstance(op, DiagonalOp):
        return DiagonalOp(np.real(op.d).astype(float))
    if isinstance(op, BandedOp):
        return BandedOp(op.n, list(op.offsets), [np.real(w).astype(float) for w in op.weights])
    return ShiftOp(op.n, op.s, float(np.real(op.factor)))


def imagpart(op):
    """``imagpart`` (src/localops.jl:157-167)."""
    if isinstance(op, DenseOp):
        return DenseOp(np.imag(op.M).astype(float))
    if isinstance(op, DiagonalOp):
        return DiagonalOp(np.imag(op.d).astype(float))
    if isinstance(op, BandedOp):
        return BandedOp(op.n, list(op.offsets), [np.imag(w).astype(float) for w in op.weights])
    return ShiftOp(op.n, op.s, float(np.imag(op.factor)))


def infnorm(op):
    """Maximum absolute row sum from the op's data (src/localops.jl:174-188)."""
    if isinstance(op, DenseOp):
        return float(np.max(np.sum(np.abs(op.M), axis=1))) if op.M.size else 0.0
    if isinstance(op, DiagonalOp):
        return float(np.max(np.abs(op.d), initial=0.0))
    if isinstance(op, BandedOp):
        m = 0.0
        for i in range(op.n):
            r = 0.0
            for o, w in zip(op.offsets, op.weights):
                if 0 <= i + o < op.n:
                    r += abs(w[i])
            m = max(m, r)
        return m
    return float(abs(op.factor)) if abs(op.s) < op.n else 0.0


# ----------------------------------------------------------------------------------------
# Bases (src/bases.jl)

_FD1 = {2: [Fr(-1, 2), 0, Fr(1, 2)],
        4: [Fr(1, 12), Fr(-2, 3), 0, Fr(2, 3), Fr(-1, 12)],
        6: [Fr(-1, 60), Fr(3, 20), Fr(-3, 4), 0, Fr(3, 4), Fr(-3, 20), Fr(1, 60)],
        8: [Fr(1, 280), Fr(-4, 105), Fr(1, 5), Fr(-4, 5), 0, Fr(4, 5), Fr(-1, 5), Fr(4, 105), Fr(-1, 280)]}


def _fd_derive(order, derivative):
    """Central FD weights solved from the Taylor moment conditions, exactly
    (the textbook construction behind src/bases.jl:124-135)."""
    h = order // 2
    xs = list(range(-h, h + 1))
    n = len(xs)
    A = [[Fr(x) ** k for x in xs] for k in range(n)]
    b = [Fr(0)] * n
    b[derivative] = Fr(math.factorial(derivative))
    # Gaussian elimination in exact arithmetic
    M = [row[:] + [b[i]] for i, row in enumerate(A)]
    for c in range(n):
        p = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[p] = M[p], M[c]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c] / M[c][c]
                M[r] = [a - f * bb for a, bb in zip(M[r], M[c])]
    return [M[i][n] / M[i][i] for i in range(n)]


def fd_weights(order, derivative):
    """``fd_weights(order, derivative)`` (src/bases.jl:124-150): the first
    derivative table is the reference's literal one; the second derivative is
    derived from the moment conditions (and the test pins both against each
    other and the reference's literal values)."""
    if order not in (2, 4, 6, 8):
        raise ValueError("order must be 2, 4, 6 or 8")
    if derivative == 1:
        return [float(x) for x in _FD1[order]]
    if derivative == 2:
        return [float(x) for x in _fd_derive(order, 2)]
    raise ValueError("derivative must be 1 or 2")


@dataclass
class ChargeBasis:
    """``ChargeBasis(ncut)``, n = -ncut..ncut (src/bases.jl:16-23)."""
    ncut: int

    @property
    def dim(self):
        return 2 * self.ncut + 1


@dataclass
class PhaseGrid:
    """``PhaseGrid(a, b, N; order=8)`` (src/bases.jl:31-43)."""
    a: float
    b: float
    N: int
    order: int = 8

    @property
    def dim(self):
        return self.N

    def coordinates(self):
        return np.linspace(self.a, self.b, self.N)

    def spacing(self):
        return (self.b - self.a) / (self.N - 1)


@dataclass
class PeriodicGrid:
    """``PeriodicGrid(N; order=8)`` on [-pi, pi) (src/bases.jl:45-91)."""
    N: int
    order: int = 8

    @property
    def dim(self):
        return self.N

    def coordinates(self):
        return -np.pi + 2 * np.pi * np.arange(self.N) / self.N

    def spacing(self):
        return 2 * np.pi / self.N


@dataclass
class FockBasis:
    """``FockBasis(N; phi_zpf)`` (src/bases.jl:93-105)."""
    N: int
    phi_zpf: float

    @property
    def dim(self):
        return self.N


def _cyclic_stencil(N, h, c):
    """src/bases.jl:215-223: every nonzero j as offset j plus its wrap j+N / j-N,
    each weight vector zero where the offset leaves 0..N-1."""
    offs, ws = [], []
    for j in range(-h, h + 1):
        if c[j + h] == 0:
            continue
        cand = (j, j + N) if j < 0 else ((j, j - N) if j > 0 else (j,))
        for o in cand:
            offs.append(o)
            ws.append(np.array([c[j + h] if 0 <= i + o < N else 0 * c[j + h] for i in range(N)]))
    return BandedOp(N, offs, ws)


def localop(b, port):
    """``localop(basis, port)`` (src/bases.jl:162-281); port = (kind, k)."""
    kind, k = port
    if isinstance(b, ChargeBasis):
        n = np.arange(-b.ncut, b.ncut + 1, dtype=float)
        if kind == "n":
            return DiagonalOp(n)
        if kind == "n2":
            return DiagonalOp(n ** 2)
        if kind == "expiphi":
            return ShiftOp(b.dim, int(round(k)), 1.0)
        if kind == "exp2pin":
            return DiagonalOp(np.exp(2j * np.pi * k * n))
    elif isinstance(b, PhaseGrid):
        N, D, h = b.N, b.spacing(), b.order // 2
        phi = b.coordinates()
        if kind == "n":
            w = fd_weights(b.order, 1)
            offs = [j for j in range(-h, h + 1) if j != 0]
            return BandedOp(N, offs, [np.full(N, -1j * w[j + h] / D) for j in offs])
        if kind == "n2":
            w = fd_weights(b.order, 2)
            return BandedOp(N, list(range(-h, h + 1)), [np.full(N, -w[j + h] / D ** 2) for j in range(-h, h + 1)])
        if kind == "phi":
            return DiagonalOp(phi)
        if kind == "phi2":
            return DiagonalOp(phi ** 2)
        if kind == "expiphi":
            return DiagonalOp(np.exp(1j * k * phi))
    elif isinstance(b, PeriodicGrid):
        N, D, h = b.N, b.spacing(), b.order // 2
        if kind == "n":
            return _cyclic_stencil(N, h, [-1j * w / D for w in fd_weights(b.order, 1)])
        if kind == "n2":
            return _cyclic_stencil(N, h, [-w / D ** 2 for w in fd_weights(b.order, 2)])
        if kind == "expiphi":
            return DiagonalOp(np.exp(1j * int(round(k)) * b.coordinates()))
    elif isinstance(b, FockBasis):
        N, pz = b.N, b.phi_zpf
        nz = 1.0 / (2 * pz)
        i1 = np.arange(1, N + 1, dtype=float)
        up = np.where(i1 < N, np.sqrt(i1), 0.0)
        dn = np.sqrt(i1 - 1)
        diag2 = 2 * (i1 - 1) + 1.0
        up2 = np.where(i1 <= N - 2, np.sqrt(i1 * (i1 + 1)), 0.0)
        dn2 = np.where(i1 >= 3, np.sqrt(np.maximum((i1 - 2) * (i1 - 1), 0)), 0.0)
        if kind == "phi":
            return BandedOp(N, [1, -1], [pz * up, pz * dn])
        if kind == "n":
            return BandedOp(N, [1, -1], [-1j * nz * up, 1j * nz * dn])
        if kind == "phi2":
            return BandedOp(N, [0, 2, -2], [pz ** 2 * diag2, pz ** 2 * up2, pz ** 2 * dn2])
        if kind == "n2":
            return BandedOp(N, [0, 2, -2], [nz ** 2 * diag2, -nz ** 2 * up2, -nz ** 2 * dn2])
    raise ValueError(f"port {port} not available on {type(b).__name__}")


def character(b, port):
    """Conjugation character of a port (src/character.jl:32-66)."""
    kind, k = port
    if isinstance(b, ChargeBasis):
        return "paired" if kind == "exp2pin" and k != 0 else "real"
    if isinstance(b, PhaseGrid):
        if kind == "n":
            return "imaginary"
        if kind == "expiphi":
            return "paired" if k != 0 else "real"
        return "real"
    if isinstance(b, PeriodicGrid):
        if kind == "n":
            return "imaginary"
        if kind == "expiphi":
            return "paired" if k != 0 else "real"
        return "real"
    if isinstance(b, FockBasis):
        if kind == "n":
            return "imaginary"
        if kind == "expiphi":
            return "paired" if k != 0 else "real"
        return "real"
    raise TypeError(b)


# ----------------------------------------------------------------------------------------
# Apply (src/apply.jl)


def _as_tensor(x, dims):
    """Column-major state tensor with a trailing block axis (src/apply.jl:18-19, 420-423)."""
    x = np.asarray(x)
    b = 1 if x.ndim == 1 else x.shape[1]
    return x.reshape(list(dims) + [b], order="F")


def apply_axis(op, X, k):
    """``apply_axis!(Y, op, X, k)`` (src/apply.jl:31-128) with X a tensor whose
    axis k (0-based) has length op.n; returns a new array."""
    Xm = np.moveaxis(X, k, 0)
    n = Xm.shape[0]
    if isinstance(op, DenseOp):
        Y = np.tensordot(op.M, Xm, axes=(1, 0))
    elif isinstance(op, DiagonalOp):
        Y = op.d.reshape([n] + [1] * (Xm.ndim - 1)) * Xm
    elif isinstance(op, BandedOp):
        dt = np.result_type(Xm.dtype, *[np.asarray(w).dtype for w in op.weights]) if op.weights else Xm.dtype
        Y = np.zeros(Xm.shape, dtype=dt)
        for o, w in zip(op.offsets, op.weights):
            lo, hi = max(0, -o), min(n, n - o)   # rows i with 0 <= i+o < n (src/apply.jl:66-79)
            if lo < hi:
                ww = np.asarray(w)[lo:hi].reshape([hi - lo] + [1] * (Xm.ndim - 1))
                Y[lo:hi] += ww * Xm[lo + o:hi + o]
    elif isinstance(op, ShiftOp):
        Y = np.zeros(Xm.shape, dtype=np.result_type(Xm.dtype, op.factor))
        s = op.s
        lo, hi = max(0, s), min(n, n + s)        # Y[t] = f X[t-s] (src/apply.jl:111-118)
        if lo < hi:
            Y[lo:hi] = op.factor * Xm[lo - s:hi - s]
    else:
        raise TypeError(op)
    return np.moveaxis(Y, 0, k)


@dataclass
class AxisTerm:
    """``AxisTerm(coeff, [(axis, op), ...])``, axes 0-based strictly increasing
    (src/apply.jl:157-166)."""
    coeff: complex
    factors: list


@dataclass
class KronOp:
    """``KronFreeOperator(dims, terms; constant)`` (src/apply.jl:222-253): all
    DiagonalOp terms and the constant fused into ``diag``; the rest kept."""
    dims: list
    diag: np.ndarray
    terms: list
    diagterms: list
    dtype: type

    @property
    def N(self):
        return int(np.prod(self.dims))


def _is_complex(v):
    return isinstance(v, complex) or np.iscomplexobj(v)


def kron_operator(dims, terms, constant=0.0):
    """Element type by promotion of the constant, coefficients and factor data
    types, as the reference does (src/apply.jl:232-233)."""
    dims = [int(d) for d in dims]
    vals = [constant] + [t.coeff for t in terms]
    for t in terms:
        for _, op in t.factors:
            vals.append(_op_sample(op))
    T = np.complex128 if any(_is_complex(v) for v in vals) else np.float64
    D = np.full(dims, constant, dtype=T, order="F")
    kept, fused = [], []
    if constant != 0:    # the constant is a fused term with no factors (src/apply.jl:237)
        fused.append(AxisTerm(constant, []))
    for t in terms:
        if not t.factors or all(isinstance(op, DiagonalOp) for _, op in t.factors):
            _add_term_diagonal(D, t.coeff, t.factors)
            fused.append(t)
        else:
            kept.append(t)
    return KronOp(dims, D.reshape(-1, order="F"), kept, fused, T)


def _op_sample(op):
    if isinstance(op, DenseOp):
        return op.M
    if isinstance(op, DiagonalOp):
        return op.d
    if isinstance(op, BandedOp):
        return np.concatenate([np.asarray(w) for w in op.weights]) if op.weights else np.zeros(0)
    return op.factor


def _add_term_diagonal(D, coeff, factors):
    """D += coeff * outer(diagonal(op_k)) (src/apply.jl:438-447)."""
    acc = np.full([1] * D.ndim, coeff, dtype=np.result_type(D.dtype, type(coeff)))
    for a, op in factors:
        shape = [1] * D.ndim
        shape[a] = D.shape[a]
        acc = acc * diagonal(op).reshape(shape)
    D += acc if np.iscomplexobj(D) else np.real(acc)


def apply_chain(H, x):
    """``mul!(y, H, x)`` (src/apply.jl:416-432): y = diag.*x, then for every
    kept term its factors chained X -> buf1 -> buf2 ... and y .+= c*src
    (_accumulate_term!, src/apply.jl:370-393)."""
    X = _as_tensor(x, H.dims)
    dt = np.result_type(H.dtype, X.dtype)
    diag = H.diag.reshape(list(H.dims) + [1], order="F")
    Y = (diag * X).astype(dt)
    for t in H.terms:
        src = X
        for a, op in t.factors:
            src = apply_axis(op, src, a)
        Y = Y + t.coeff * src
    out = Y.reshape(-1, X.shape[-1], order="F")
    return out[:, 0] if np.asarray(x).ndim == 1 else out


def kron_reference(dims, terms, constant=0.0):
    """Dense sum_t c_t kron(A_K, ..., A_1) with identities (test/helpers.jl:10-24);
    terms as AxisTerm with ops (dense() taken here)."""
    D = int(np.prod(dims))
    Hm = np.zeros((D, D), dtype=complex)
    for t in terms:
        ops = {a: dense(op) for a, op in t.factors}
        T = np.ones((1, 1), dtype=complex)
        for a in range(len(dims)):
            T = np.kron(ops.get(a, np.eye(dims[a])), T)
        Hm += t.coeff * T
    return Hm + constant * np.eye(D)


def to_matrix(H):
    """``Matrix(H)`` by applying H to identity columns (src/apply.jl:506-531)."""
    N = H.N
    return apply_chain(H, np.eye(N, dtype=H.dtype))


def full_diagonal(H):
    """``_full_diagonal(H)`` (src/apply.jl:459-478)."""
    D = H.diag.reshape(H.dims, order="F").astype(complex)
    for t in H.terms:
        acc = np.full([1] * len(H.dims), t.coeff, dtype=complex)
        for a, op in t.factors:
            shape = [1] * len(H.dims)
            shape[a] = H.dims[a]
            acc = acc * diagonal(op).reshape(shape)
        D = D + acc
    return np.real(D.reshape(-1, order="F"))


def gershgorin_bound(H):
    """``gershgorin_bound(H)`` (src/apply.jl:586-593)."""
    off = sum(abs(t.coeff) * np.prod([infnorm(op) for _, op in t.factors]) for t in H.terms)
    return float(np.max(np.real(H.diag)) + np.max(np.abs(np.imag(H.diag))) + off)


def separable_eigenbases(H):
    """``separable_eigenbases(H)`` (src/apply.jl:606-642)."""
    T = complex if H.dtype == np.complex128 else float
    A = [np.zeros((d, d), dtype=T) for d in H.dims]
    for t in H.terms:
        if len(t.factors) == 1:
            a, op = t.factors[0]
            A[a] = A[a] + t.coeff * dense(op)
    for t in H.diagterms:
        if len(t.factors) == 1:
            a, op = t.factors[0]
            A[a] = A[a] + t.coeff * dense(op)
        elif len(t.factors) == 0:
            A[0] = A[0] + t.coeff * np.eye(H.dims[0])
    U, lam = [], []
    for k, M in enumerate(A):
        asym, nrm = np.linalg.norm(M - M.conj().T), np.linalg.norm(M)
        if not asym <= math.sqrt(np.finfo(float).eps) * nrm:
            raise ValueError(f"separable_eigenbases: the single-axis terms on axis {k} sum to a non-Hermitian "
                             f"matrix (|M - M'| = {asym}, |M| = {nrm}), so the operator is not Hermitian")
        if np.count_nonzero(M - np.diag(np.diag(M))) == 0:
            U.append(np.eye(M.shape[0], dtype=T))
            lam.append(np.real(np.diag(M)).copy())
        else:
            w, V = np.linalg.eigh((M + M.conj().T) / 2)
            U.append(V)
            lam.append(w)
    return U, lam


def product_seed(H, k):
    """``product_seed(H, k)`` (src/apply.jl:655-685), ties broken by the first
    (stable) index as partialsortperm does."""
    U, lam = separable_eigenbases(H)
    N = H.N
    X = np.zeros((N, k), dtype=H.dtype)
    if all(np.array_equal(u, np.eye(u.shape[0])) for u in U):
        order = np.argsort(full_diagonal(H), kind="stable")[:k]
        for j, i in enumerate(order):
            X[i, j] = 1
        return X
    sums = lam[0]
    for a in range(1, len(H.dims)):
        sums = (sums[:, None] + lam[a][None, :]).reshape(-1, order="F")
    order = np.argsort(sums, kind="stable")[:k]
    for j, flat in enumerate(order):
        idx = np.unravel_index(flat, H.dims, order="F")
        v = np.ones(1, dtype=H.dtype)
        for a in range(len(H.dims)):
            v = np.kron(U[a][:, idx[a]], v)
        X[:, j] = v
    return X


# ----------------------------------------------------------------------------------------
# Circuit terms and the real node expansion (src/terms.jl:119-168, src/character.jl:276-315)


@dataclass
class NumericSpec:
    """A kronfree spec with its coefficients evaluated at theta
    (src/spec.jl:33-52 canonical form)."""
    modes: list
    E_C: np.ndarray
    E_L: np.ndarray
    flux_linear: np.ndarray
    n_g: np.ndarray
    cosines: list                      # (energy, omega list, offset)
    constant: float = 0.0


def circuit_terms(spec):
    """``circuit_terms`` (src/terms.jl:119-157): list of (coeff, [(mode, port)])."""
    K = len(spec.modes)
    E_C, E_L = np.asarray(spec.E_C, float), np.asarray(spec.E_L, float)
    out = []

    def add(c, fs):
        if c != 0:
            out.append((c, fs))
    for i in range(K):
        add(4 * E_C[i, i], [(i, ("n2", 0.0))])
    for i in range(K):
        for j in range(i + 1, K):
            add(4 * (E_C[i, j] + E_C[j, i]), [(i, ("n", 0.0)), (j, ("n", 0.0))])
    for i in range(K):
        lin = sum((E_C[i, j] + E_C[j, i]) * spec.n_g[j] for j in range(K))
        add(-4 * lin, [(i, ("n", 0.0))])
    for i in range(K):
        add(0.5 * E_L[i, i], [(i, ("phi2", 0.0))])
    for i in range(K):
        for j in range(i + 1, K):
            add(0.5 * (E_L[i, j] + E_L[j, i]), [(i, ("phi", 0.0)), (j, ("phi", 0.0))])
    for i in range(K):
        add(spec.flux_linear[i], [(i, ("phi", 0.0))])
    for energy, omega, offset in spec.cosines:
        for s in (1, -1):
            fs = [(j, ("expiphi", float(s * omega[j]))) for j in range(K) if omega[j] != 0]
            add(-0.5 * energy * np.exp(1j * s * offset), fs)
    return out


def constant_term(spec):
    """``constant_term`` (src/terms.jl:164-168)."""
    ng = np.asarray(spec.n_g, float)
    return spec.constant + 4 * float(ng @ np.asarray(spec.E_C, float) @ ng)


def _re_ipow(c, m):
    """Re(c * i^m) exactly (src/character.jl:282-285)."""
    r = m % 4
    c = complex(c)
    return [c.real, -c.imag, -c.real, c.imag][r]


def real_expansion(c, factors):
    """``_real_expansion`` on the factors' own components (src/character.jl:297-315):
    factors = [(axis, op, kind)], returns [AxisTerm] of Re(c x O), unmerged."""
    comps = []
    for a, op, kind in factors:
        R = None if kind == "imaginary" else realpart(op)
        I = None if kind == "real" else imagpart(op)
        ch = (False,) if kind == "real" else ((True,) if kind == "imaginary" else (False, True))
        comps.append((a, R, I, ch))
    out = []
    for S in itertools.product(*[x[3] for x in comps]):
        coeff = _re_ipow(c, sum(S))
        if coeff == 0:
            continue
        ops = [(x[0], x[2] if inS else x[1]) for x, inS in zip(comps, S)]
        out.append(AxisTerm(coeff, ops))
    return out


# ----------------------------------------------------------------------------------------
# A small tree: leaves of bases, branches of nodes; Dense solves (src/materialize.jl)


@dataclass
class Node:
    """Node(children; keep) (src/tree.jl:419-456): children are (mode_name, basis)
    pairs or Nodes; keep = number of lowest states kept."""
    children: list
    keep: int
    real: bool = True
    solver: object = None   # None: dense (numpy eigh); callable(H, m) -> (E, V) otherwise

    def modes(self):
        out = []
        for c in self.children:
            out += [c[0]] if isinstance(c, tuple) else c.modes()
        return out


@dataclass
class Materialized:
    node: Node
    dims: list
    energies: np.ndarray
    vectors: np.ndarray
    children: list
    H: KronOp = None
    all_energies: np.ndarray = None


def _route(root, terms):
    """``route`` (src/tree.jl:500-529): each term goes to the lowest node whose
    modes contain all of the term's modes."""
    routed = {}

    def lowest(nd, modes):
        for ch in nd.children:
            if isinstance(ch, Node) and all(m in ch.modes() for m in modes):
                return lowest(ch, modes)
        return nd
    for t in terms:
        nd = lowest(root, [f[0] for f in t[1]])
        routed.setdefault(id(nd), []).append(t)
    return routed


def _leaf_basis(node, mode):
    for c in node.children:
        if isinstance(c, tuple):
            if c[0] == mode:
                return c[1]
        elif mode in c.modes():
            return _leaf_basis(c, mode)
    raise KeyError(mode)


def _child_index(node, mode):
    for a, c in enumerate(node.children):
        if (isinstance(c, tuple) and c[0] == mode) or (isinstance(c, Node) and mode in c.modes()):
            return a
    raise KeyError(mode)


def _mode_op(mat, mode, port):
    """The operator of ``port`` on ``mode`` in the product basis of ``mat``'s
    node, applied on its axis (a basis child's localop or a node child's port)."""
    node = mat.node
    a = _child_index(node, mode)
    ch = node.children[a]
    if isinstance(ch, tuple):
        return a, localop(ch[1], port)
    return a, DenseOp(port_matrix(mat.children[a], mode, port))


def port_matrix(mat, mode, port):
    """``_port_matrix`` (src/materialize.jl:443-459): V' (O on mode's axis) V."""
    a, op = _mode_op(mat, mode, port)
    V = mat.vectors
    T = _as_tensor(V, mat.dims)
    OV = apply_axis(op, T, a).reshape(V.shape, order="F")
    return V.conj().T @ OV


def joint_port(mat, factors):
    """``_joint_port`` (src/materialize.jl:490-518): V' (x_a O_a) V."""
    V = mat.vectors
    T = _as_tensor(V, mat.dims)
    for mode, port in factors:
        a, op = _mode_op(mat, mode, port)
        T = apply_axis(op, T, a)
    return V.conj().T @ T.reshape(V.shape, order="F")


def node_terms(node, children_mats, terms, real):
    """The AxisTerms of a node (``_node_operator``, src/materialize.jl:583-603):
    routed terms on the children's operators (real expansion on a real node),
    plus each node child's kept energies as a DiagonalOp on its axis."""
    axisterms = []
    for c, fs in terms:
        groups = {}
        for mode, port in fs:
            groups.setdefault(_child_index(node, mode), []).append((mode, port))
        ops = []
        for a in sorted(groups):
            g = groups[a]
            ch = node.children[a]
            if len(g) == 1:
                mode, port = g[0]
                if isinstance(ch, tuple):
                    op, kind = localop(ch[1], port), character(ch[1], port)
                else:
                    op = DenseOp(port_matrix(children_mats[a], mode, port))
                    kind = character(_leaf_basis(node, mode), port)
            else:
                op = DenseOp(joint_port(children_mats[a], g))
                kinds = [character(_leaf_basis(node, m), p) for m, p in g]
                if "paired" in kinds:
                    kind = "paired"
                else:
                    kind = "imaginary" if kinds.count("imaginary") % 2 else "real"
            ops.append((a, op, kind))
        if real:
            axisterms += real_expansion(c, ops)
        else:
            axisterms.append(AxisTerm(complex(c), [(a, op) for a, op, _ in ops]))
    for a, ch in enumerate(node.children):
        if isinstance(ch, Node):
            axisterms.append(AxisTerm(1.0, [(a, DiagonalOp(np.asarray(children_mats[a].energies, float)))]))
    return axisterms


def dense_solve(H, m):
    """``solve(::Dense, H, m)`` (src/solvers.jl:44-50) with numpy's LAPACK eigh."""
    M = to_matrix(H)
    M = (M + M.conj().T) / 2
    w, V = np.linalg.eigh(M)
    if 0 < m < len(w):
        return w[:m], V[:, :m], w
    return w, V, w


def materialize(root, spec, solver=None):
    """``materialize(root, spec, theta)`` (src/materialize.jl:177-192, 713-731)
    with Dense solves (or ``node.solver``); returns the root Materialized."""
    terms = circuit_terms(spec)
    names = spec.modes
    named = [(c, [(names[m], p) for m, p in fs]) for c, fs in terms]
    routed = _route(root, named)
    const = constant_term(spec)

    def visit(nd, is_root):
        kids = []
        for ch in nd.children:
            kids.append(None if isinstance(ch, tuple) else visit(ch, False))
        dims = [ch[1].dim if isinstance(ch, tuple) else len(k.energies) for ch, k in zip(nd.children, kids)]
        ts = node_terms(nd, kids, routed.get(id(nd), []), nd.real)
        H = kron_operator(dims, ts, const if is_root else 0.0)
        solve = nd.solver or (lambda H_, m_: dense_solve(H_, m_))
        E, V, allE = solve(H, nd.keep)
        return Materialized(nd, dims, np.asarray(E)[:nd.keep], np.asarray(V)[:, :nd.keep], kids, H, allE)
    return visit(root, True)


# ----------------------------------------------------------------------------------------
# Specs of the reference fixtures and of the benchmark configurations


def transmon_spec(E_J, E_C, n_g):
    """test/fixtures/transmon.json."""
    return NumericSpec(["varphi_1"], [[E_C]], [[0.0]], [0.0], [n_g], [(E_J, [1], 0.0)])


def fluxonium_spec(E_J, E_C, E_L, varphi_ext):
    """test/fixtures/fluxonium.json."""
    return NumericSpec(["varphi_1"], [[E_C]], [[E_L]], [E_L * varphi_ext], [0.0], [(E_J, [1], 0.0)],
                       E_L * varphi_ext ** 2 / 2)


def two_transmons_spec(E_J1, E_C1, n_g1, E_J2, E_C2, n_g2, g):
    """test/fixtures/two_transmons.json."""
    return NumericSpec(["varphi_1", "varphi_2"], [[E_C1, g / 8], [g / 8, E_C2]], np.zeros((2, 2)), [0.0, 0.0],
                       [n_g1, n_g2], [(E_J1, [1, 0], 0.0), (E_J2, [0, 1], 0.0)])


def zero_pi_spec(E_J, E_L, E_CJ, E_CS, varphi_ext, n_g):
    """test/fixtures/zero_pi.json."""
    x = varphi_ext / 2
    return NumericSpec(["theta", "phi"], [[E_CS / 2, 0.0], [0.0, E_CJ / 2]], [[0.0, 0.0], [0.0, 2 * E_L]],
                       [0.0, 0.0], [n_g, 0.0], [(E_J, [1, 1], -x), (E_J, [1, -1], x)], 2 * E_J)


def full_zero_pi_spec(E_J, E_L, E_CJ, E_CS, E_Cz, dE_J, dC_J, dC, dE_L, varphi_ext, n_g):
    """test/fixtures/full_zero_pi.json."""
    s2 = math.sqrt(2.0)
    D = 1 - E_CS * dC_J ** 2 / (4 * E_CJ) - E_CS * dC ** 2 / (4 * E_Cz)
    E_C = [[E_CS / 2, E_CS * dC_J / 4, -s2 * E_CS * dC / 4],
           [E_CS * dC_J / 4, E_CJ / 2, 0.0],
           [-s2 * E_CS * dC / 4, 0.0, E_Cz]]
    E_Lm = [[0.0, 0.0, 0.0], [0.0, 2 * E_L, E_L * dE_L / s2], [0.0, E_L * dE_L / s2, E_L]]
    ng = [n_g / D, -n_g * E_CS * dC_J / (2 * E_CJ) / D, n_g * s2 * E_CS * dC / (4 * E_Cz) / D]
    x = varphi_ext / 2
    cos = [(E_J, [1, 1, 0], -x), (E_J, [1, -1, 0], x), (E_J * dE_J / 2, [1, 1, 0], -x), (-E_J * dE_J / 2, [1, -1, 0], x)]
    const = 2 * E_J - math.sqrt(2 * E_Cz * E_L) + 2 * E_CS * n_g ** 2 - 2 * E_CS * n_g ** 2 / D
    return NumericSpec(["theta", "phi", "zeta"], E_C, E_Lm, [0.0, 0.0, 0.0], ng, cos, const)


def three_fluxonia_spec(EC, EL, EJ, phiext, g=0.05):
    """Three capacitively coupled fluxonia: per mode 4 E_C n^2 + E_L phi^2/2 -
    E_J cos(phi - phi_ext), coupling g n_i n_j for each pair (E_C[i,j] = g/8)."""
    E_C = np.full((3, 3), g / 8)
    for i in range(3):
        E_C[i, i] = EC[i]
    return NumericSpec(["phi1", "phi2", "phi3"], E_C, np.diag(EL), [0.0] * 3, [0.0] * 3,
                       [(EJ[i], [int(i == j) for j in range(3)], -phiext[i]) for i in range(3)])


def three_fluxonia_operator(EC, EL, EJ, phiext, N=512, span=6 * np.pi, order=2, g=0.05, real=True):
    """The flat three-fluxonia leaf operator (one node of three PhaseGrid modes)."""
    spec = three_fluxonia_spec(EC, EL, EJ, phiext, g)
    node = Node([(m, PhaseGrid(-span, span, N, order)) for m in spec.modes], keep=1, real=real)
    terms = circuit_terms(spec)
    named = [(c, [(spec.modes[m], p) for m, p in fs]) for c, fs in terms]
    ts = node_terms(node, [None] * 3, named, real)
    return kron_operator([N] * 3, ts, constant_term(spec))


# ----------------------------------------------------------------------------------------
# Iterative solver restatements on the oracle apply (for residual checks and
# small-size behaviour); Dense/eigh pins their results in the tests.


def true_residuals(H, E, V):
    """``_worst_true_residual`` columns (src/solvers.jl:65-69): ||H v - E v||."""
    R = apply_chain(H, V) - V * np.asarray(E)[None, :]
    return np.linalg.norm(R, axis=0)


def residual_floor(hbound, E1):
    """``_residual_floor`` (src/solvers.jl:75)."""
    return 10 * np.finfo(float).eps * max(abs(hbound), abs(E1))


# ----------------------------------------------------------------------------------------
# The linear solve on the complement of a node's known states (the Sternheimer step of
# the derivatives), src/resolvent.jl:119-256.  Pinned in tests/test_oracle.py against the
# spectral solution x_j = sum over the states outside the basis of
# psi_l <psi_l|Q b_j> / (E_l - s_j) on dense matrices, and against its own stopping rules.

FLOOR_CAP = 1000       # src/resolvent.jl:129 (_FLOOR_CAP), units of eps |R_j|
FLOOR_PATIENCE = 5     # src/resolvent.jl:131 (_FLOOR_PATIENCE)


class ComplementError(Exception):
    """The solve's ErrorException (src/resolvent.jl:199-203, 235-239, 246-252)."""


def _colredot(A, B):
    """Re<A_j, B_j> per column (src/resolvent.jl:124)."""
    return np.real(np.sum(np.conj(A) * B, axis=0))


def complement_solve(apply, D, basis, B, shifts, tol, maxiter, eta=0.01, scale=None):
    """``_complement_solve`` (src/resolvent.jl:146-256): X[:, j] solves
    Q (H - s_j) Q x = Q B[:, j] with x orthogonal to ``basis`` (Q = 1 - basis basis'),
    by preconditioned CG on all columns in lockstep, each with its own scalars,
    preconditioned by 1 / hypot(D - s_j, eta).  ``apply(X)`` is H X.  Returns
    (X, iterations, residuals, from_start, at_floor, floor_residuals), indices 0-based."""
    B = np.asarray(B)
    n = B.shape[1]
    shifts = np.asarray(shifts, dtype=float)
    if len(shifts) != n:
        raise ValueError(f"{n} columns and {len(shifts)} shifts")
    scale = np.linalg.norm(B, axis=0) if scale is None else np.asarray(scale, dtype=float)
    Phi = np.asarray(basis)

    def project(Y):
        return Y - Phi @ (Phi.conj().T @ Y)

    S = shifts[None, :]
    Minv = 1.0 / np.hypot(np.asarray(D)[:, None] - S, eta)
    cap = FLOOR_CAP * np.finfo(float).eps * scale
    QB = project(B.copy())
    b = np.linalg.norm(QB, axis=0)
    goal = tol * b
    from_start = [j for j in range(n) if b[j] <= cap[j]]
    met = np.zeros(n, dtype=bool)
    atfloor = np.zeros(n, dtype=bool)
    atfloor[from_start] = True
    QB = QB * (~atfloor)[None, :]
    X = np.zeros_like(QB)
    R = QB.copy()
    Z = project(Minv * R)
    P = Z.copy()
    rho = _colredot(R, Z)

    def true_residuals():
        T = apply(X) - X * S
        T = project(T) - QB
        return np.linalg.norm(T, axis=0)

    rn = b.copy()
    anchor = np.full(n, np.inf)
    stalled = np.zeros(n, dtype=int)
    it = 0
    while not np.all(met | atfloor) and it < maxiter:
        it += 1
        AP = project(apply(P) - P * S)
        kappa = _colredot(P, AP)
        for j in range(n):
            if met[j] or atfloor[j] or kappa[j] > 0:
                continue
            if rn[j] <= cap[j]:
                atfloor[j] = True
                continue
            raise ComplementError(f"the linear solve on the complement is not positive on the complement at column {j + 1}")
        active = ~(met | atfloor)
        alpha = np.where(active, rho / np.where(active, kappa, 1.0), 0.0)
        X = X + P * alpha[None, :]
        R = R - AP * alpha[None, :]
        rn = np.linalg.norm(R, axis=0)
        met |= active & (rn <= goal)
        near = ~(met | atfloor) & (rn <= cap)
        if np.any(near):
            tr = true_residuals()
            for j in np.flatnonzero(near):
                if tr[j] < anchor[j] / 2:
                    anchor[j], stalled[j] = tr[j], 0
                else:
                    stalled[j] += 1
                if stalled[j] >= FLOOR_PATIENCE and tr[j] <= cap[j]:
                    atfloor[j] = True
        if np.all(met | atfloor):
            break
        Z = project(Minv * R)
        rhon = _colredot(R, Z)
        with np.errstate(divide="ignore", invalid="ignore"):
            beta = np.where(met | atfloor, 0.0, rhon / rho)
        P = Z + P * beta[None, :]
        rho = rhon
    tr = true_residuals()
    for j in range(n):
        if not (met[j] or atfloor[j]) and tr[j] <= cap[j]:
            atfloor[j] = True
    if not np.all(met | atfloor):
        j = int(np.flatnonzero(~(met | atfloor))[0])
        raise ComplementError(f"the linear solve on the complement did not converge at column {j + 1}")
    met0 = met.copy()
    for j in range(n):
        if j in from_start:
            continue
        if met[j] and tr[j] > 10 * goal[j]:
            met[j] = False
            atfloor[j] = True
        if atfloor[j] and tr[j] > cap[j]:
            raise ComplementError(f"the linear solve on the complement did not converge at column {j + 1}"
                                  + (" (met by the recurrence)" if met0[j] else ""))
    residuals = np.array([tr[j] / b[j] if met[j] else 0.0 for j in range(n)])
    at_floor = [j for j in range(n) if atfloor[j] and j not in from_start]
    floor_residuals = np.array([tr[j] / scale[j] if j in at_floor else 0.0 for j in range(n)])
    return X, it, residuals, from_start, at_floor, floor_residuals


# ----------------------------------------------------------------------------------------
# The warm-start network's FP64 oracle (no reference counterpart: the model
# maps circuit parameters to an initial LOBPCG block).


def gelu_tanh(x):
    c = math.sqrt(2.0 / math.pi)
    return 0.5 * x * (1.0 + np.tanh(c * (x + 0.044715 * x ** 3)))


def gelu_tanh_grad(x):
    c = math.sqrt(2.0 / math.pi)
    u = c * (x + 0.044715 * x ** 3)
    t = np.tanh(u)
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * c * (1 + 3 * 0.044715 * x * x)


def mlp_forward(params, X):
    """y = W3 gelu(W2 gelu(W1 x + b1) + b2) + b3, batch rows of X (FP64)."""
    W1, b1, W2, b2, W3, b3 = params
    z1 = X @ W1.T + b1
    a1 = gelu_tanh(z1)
    z2 = a1 @ W2.T + b2
    a2 = gelu_tanh(z2)
    y = a2 @ W3.T + b3
    return y, (X, z1, a1, z2, a2)


def mlp_backward(params, cache, dy):
    """Gradients of sum(dy * y) w.r.t. every parameter (FP64)."""
    W1, b1, W2, b2, W3, b3 = params
    X, z1, a1, z2, a2 = cache
    dW3 = dy.T @ a2
    db3 = dy.sum(0)
    da2 = dy @ W3
    dz2 = da2 * gelu_tanh_grad(z2)
    dW2 = dz2.T @ a1
    db2 = dz2.sum(0)
    da1 = dz2 @ W2
    dz1 = da1 * gelu_tanh_grad(z1)
    dW1 = dz1.T @ X
    db1 = dz1.sum(0)
    return dW1, db1, dW2, db2, dW3, db3
