"""Single-mode bases and their port operators (src/bases.jl)."""
from __future__ import annotations

import math

import numpy as np

from ._lib import ArgumentError
from .localops import BandedOp, DenseOp, DiagonalOp, ShiftOp

# Central finite-difference weights (src/bases.jl:124-135), exact rationals as floats.
_FD1 = {2: [-1 / 2, 0, 1 / 2],
        4: [1 / 12, -2 / 3, 0, 2 / 3, -1 / 12],
        6: [-1 / 60, 3 / 20, -3 / 4, 0, 3 / 4, -3 / 20, 1 / 60],
        8: [1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280]}
_FD2 = {2: [1.0, -2.0, 1.0],
        4: [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12],
        6: [1 / 90, -3 / 20, 3 / 2, -49 / 18, 3 / 2, -3 / 20, 1 / 90],
        8: [-1 / 560, 8 / 315, -1 / 5, 8 / 5, -205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560]}


def fd_weights(order, derivative):
    """``fd_weights(order, derivative)`` (src/bases.jl:124-150)."""
    if order not in (2, 4, 6, 8):
        raise ArgumentError(f"fd_weights: order must be 2, 4, 6 or 8, got {order}")
    if derivative == 1:
        return list(_FD1[order])
    if derivative == 2:
        return list(_FD2[order])
    raise ArgumentError(f"fd_weights: derivative must be 1 or 2, got {derivative}")


class Basis:
    pass


class ChargeBasis(Basis):
    """``ChargeBasis(ncut)``: |n>, n = -ncut..ncut (src/bases.jl:16-23)."""

    def __init__(self, ncut):
        if ncut < 0:
            raise ArgumentError(f"ChargeBasis requires ncut >= 0, got {ncut}")
        self.ncut = int(ncut)

    @property
    def dim(self):
        return 2 * self.ncut + 1


class PhaseGrid(Basis):
    """``PhaseGrid(a, b, N; order=8)`` (src/bases.jl:31-43)."""

    def __init__(self, a, b, N, order=8):
        if order not in (2, 4, 6, 8):
            raise ArgumentError(f"PhaseGrid order must be 2, 4, 6 or 8, got {order}")
        if not b > a:
            raise ArgumentError(f"PhaseGrid requires b > a, got a = {a}, b = {b}")
        if not N > order:
            raise ArgumentError(f"PhaseGrid requires N > order = {order}, got N = {N}")
        self.a, self.b, self.N, self.order = float(a), float(b), int(N), int(order)

    @property
    def dim(self):
        return self.N

    def coordinates(self):
        return np.linspace(self.a, self.b, self.N)

    def spacing(self):
        return (self.b - self.a) / (self.N - 1)


class PeriodicGrid(Basis):
    """``PeriodicGrid(N; order=8)`` on [-pi, pi) without the endpoint (src/bases.jl:45-91)."""

    def __init__(self, N, order=8):
        if order not in (2, 4, 6, 8):
            raise ArgumentError(f"PeriodicGrid order must be 2, 4, 6 or 8, got {order}")
        if N < order + 1:
            raise ArgumentError(f"PeriodicGrid requires N >= order + 1 = {order + 1}, got N = {N}")
        self.N, self.order = int(N), int(order)

    @property
    def dim(self):
        return self.N

    def coordinates(self):
        return -np.pi + 2 * np.pi * np.arange(self.N) / self.N

    def spacing(self):
        return 2 * np.pi / self.N


class FockBasis(Basis):
    """``FockBasis(N; phi_zpf)`` (src/bases.jl:93-105)."""

    def __init__(self, N, phi_zpf):
        if N < 1:
            raise ArgumentError(f"FockBasis requires N >= 1, got {N}")
        if not phi_zpf > 0:
            raise ArgumentError(f"FockBasis requires phi_zpf > 0, got {phi_zpf}")
        self.N, self.phi_zpf = int(N), float(phi_zpf)

    @property
    def dim(self):
        return self.N


def cyclic_stencil(N, h, c):
    """``_cyclic_stencil`` (src/bases.jl:215-223): every nonzero j as offset j
    plus its wrap j+N (j<0) or j-N (j>0), each weight vector zero where the
    offset leaves the grid."""
    offs, ws = [], []
    i = np.arange(N)
    for j in range(-h, h + 1):
        cj = c[j + h]
        if cj == 0:
            continue
        for o in ((j, j + N) if j < 0 else ((j, j - N) if j > 0 else (j,))):
            offs.append(o)
            ws.append(np.where((i + o >= 0) & (i + o < N), cj, 0 * cj))
    return BandedOp(N, offs, ws)


_LOCALOP_MEMO = {}


def localop(b, port):
    """``localop(basis, port)`` (src/bases.jl:162-281); port = (kind, k).  The
    operators are read-only and a function of (basis, port) alone: a sweep's many
    nodes on the same few grids share them (up to 1024 remembered)."""
    key = (type(b).__name__, tuple(sorted(vars(b).items())), port[0], port[1])
    op = _LOCALOP_MEMO.get(key)
    if op is None:
        op = _localop(b, port)
        if len(_LOCALOP_MEMO) >= 1024:
            _LOCALOP_MEMO.clear()
        _LOCALOP_MEMO[key] = op
    return op


def _localop(b, port):
    kind, k = port
    if isinstance(b, ChargeBasis):
        n = np.arange(-b.ncut, b.ncut + 1, dtype=float)
        if kind == "n":
            return DiagonalOp(n)
        if kind == "n2":
            return DiagonalOp(n ** 2)
        if kind == "expiphi":
            if abs(k - round(k)) > 1e-12:
                raise ArgumentError(f"ChargeBasis: port expiphi({k}) needs an integer k")
            return ShiftOp(b.dim, int(round(k)), 1.0)
        if kind == "exp2pin":
            return DiagonalOp(np.exp(2j * np.pi * k * n))
    elif isinstance(b, PhaseGrid):
        N, D, h = b.N, b.spacing(), b.order // 2
        if kind == "n":
            w = fd_weights(b.order, 1)
            offs = [j for j in range(-h, h + 1) if j != 0]
            return BandedOp(N, offs, [np.full(N, -1j * w[j + h] / D) for j in offs])
        if kind == "n2":
            w = fd_weights(b.order, 2)
            return BandedOp(N, list(range(-h, h + 1)), [np.full(N, -w[j + h] / D ** 2) for j in range(-h, h + 1)])
        if kind == "phi":
            return DiagonalOp(b.coordinates())
        if kind == "phi2":
            return DiagonalOp(b.coordinates() ** 2)
        if kind == "expiphi":
            return DiagonalOp(np.exp(1j * k * b.coordinates()))
        if kind == "exp2pin":
            s = 2 * np.pi * k / D
            if abs(s - round(s)) >= 1e-8:
                raise ArgumentError(f"PhaseGrid: port exp2pin({k}) needs 2 pi k / Delta to be an integer, got {s}")
            return ShiftOp(N, -int(round(s)), 1.0)
    elif isinstance(b, PeriodicGrid):
        N, D, h = b.N, b.spacing(), b.order // 2
        if kind == "n":
            return cyclic_stencil(N, h, [-1j * w / D for w in fd_weights(b.order, 1)])
        if kind == "n2":
            return cyclic_stencil(N, h, [-w / D ** 2 for w in fd_weights(b.order, 2)])
        if kind == "expiphi":
            if abs(k - round(k)) > 1e-12:
                raise ArgumentError(f"PeriodicGrid: port expiphi({k}) needs an integer k on a 2pi-periodic grid")
            return DiagonalOp(np.exp(1j * int(round(k)) * b.coordinates()))
    elif isinstance(b, FockBasis):
        N, pz = b.N, b.phi_zpf
        nz = 1.0 / (2 * pz)
        i1 = np.arange(1, N + 1, dtype=float)
        up = np.where(i1 < N, np.sqrt(i1), 0.0)
        dn = np.sqrt(i1 - 1)
        diag2 = 2 * (i1 - 1) + 1.0
        up2 = np.where(i1 <= N - 2, np.sqrt(i1 * (i1 + 1)), 0.0)
        dn2 = np.where(i1 >= 3, np.sqrt(np.maximum((i1 - 2) * (i1 - 1), 0.0)), 0.0)
        if kind == "phi":
            return BandedOp(N, [1, -1], [pz * up, pz * dn])
        if kind == "n":
            return BandedOp(N, [1, -1], [-1j * nz * up, 1j * nz * dn])
        if kind == "phi2":
            return BandedOp(N, [0, 2, -2], [pz ** 2 * diag2, pz ** 2 * up2, pz ** 2 * dn2])
        if kind == "n2":
            return BandedOp(N, [0, 2, -2], [nz ** 2 * diag2, -nz ** 2 * up2, -nz ** 2 * dn2])
    raise ArgumentError(f"{type(b).__name__}: port {kind}({k}) is not available")


def character(b, port):
    """Conjugation character of a port: "real", "imaginary" or "paired"
    (src/character.jl:32-66)."""
    kind, k = port
    if isinstance(b, ChargeBasis):
        return "paired" if kind == "exp2pin" and k != 0 else "real"
    if kind == "n":
        return "imaginary"
    if kind == "expiphi":
        return "paired" if k != 0 else "real"
    return "real"


def canonical_port(port, kind):
    """``_canonical_port`` (src/character.jl:276-279): the k > 0 member of a
    paired orbit and the sign of the imaginary component."""
    if kind != "paired":
        return port, 1
    name, k = port
    return ((name, k), 1) if k > 0 else ((name, -k), -1)


def positive_sqrt4(x):
    return math.sqrt(math.sqrt(x))
