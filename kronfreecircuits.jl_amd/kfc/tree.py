"""The caller side of the hot path: circuit terms, the real node operator and a
tree materialization that hands each node's KronFreeOperator to a Solver.

Follows src/terms.jl:119-168 (circuit_terms, constant_term),
src/character.jl:276-315 (_real_expansion), src/materialize.jl:583-647
(_node_operator, _real_node_terms), src/tree.jl:500-529 (route) and
src/materialize.jl:427-459 (ports V' O V), with spec coefficients given as
numbers (a spec evaluated at theta)."""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .bases import Basis, canonical_port, character, localop
from .localops import DenseOp, DiagonalOp, dense, imagpart, realpart
from .operator import AxisTerm, KronFreeOperator
from .solvers import Auto, Dense, LOBPCG, resolve, solve_stats


@dataclass
class NumericSpec:
    """The canonical form 4(n-n_g)^T E_C (n-n_g) + 1/2 phi^T E_L phi + flux_linear . phi
    - sum E cos(omega . phi + offset) + constant (src/spec.jl:33-52) at one theta."""
    modes: list
    E_C: np.ndarray
    E_L: np.ndarray
    flux_linear: list
    n_g: list
    cosines: list
    constant: float = 0.0


def circuit_terms(spec):
    """``circuit_terms(spec)`` (src/terms.jl:119-157): [(coeff, [(mode, port)])]."""
    K = len(spec.modes)
    E_C, E_L, names = np.asarray(spec.E_C, float), np.asarray(spec.E_L, float), spec.modes
    out = []

    def add(c, fs):
        if c != 0:
            out.append((c, [(names[m], p) for m, p in fs]))
    for i in range(K):
        add(4 * E_C[i, i], [(i, ("n2", 0.0))])
    for i in range(K):
        for j in range(i + 1, K):
            add(4 * (E_C[i, j] + E_C[j, i]), [(i, ("n", 0.0)), (j, ("n", 0.0))])
    for i in range(K):
        add(-4 * sum((E_C[i, j] + E_C[j, i]) * spec.n_g[j] for j in range(K)), [(i, ("n", 0.0))])
    for i in range(K):
        add(0.5 * E_L[i, i], [(i, ("phi2", 0.0))])
    for i in range(K):
        for j in range(i + 1, K):
            add(0.5 * (E_L[i, j] + E_L[j, i]), [(i, ("phi", 0.0)), (j, ("phi", 0.0))])
    for i in range(K):
        add(spec.flux_linear[i], [(i, ("phi", 0.0))])
    for energy, omega, offset in spec.cosines:
        for s in (1, -1):
            add(-0.5 * energy * np.exp(1j * s * offset),
                [(j, ("expiphi", float(s * omega[j]))) for j in range(K) if omega[j] != 0])
    return out


def constant_term(spec):
    """``constant_term`` (src/terms.jl:164-168): constant + 4 n_g^T E_C n_g."""
    ng = np.asarray(spec.n_g, float)
    return spec.constant + 4 * float(ng @ np.asarray(spec.E_C, float) @ ng)


def _re_ipow(c, m):
    c = complex(c)
    return [c.real, -c.imag, -c.real, c.imag][m % 4]


class Node:
    """``Node(children; solver=Auto(), keep=lowest(k), compute, real)`` (src/tree.jl:419-456):
    children are ``(mode, basis)`` pairs or Nodes."""

    def __init__(self, children, keep, solver=None, compute=None, real=True, ports=()):
        self.children = list(children)
        self.keep = int(keep)
        self.solver = solver or Auto()
        self.compute = int(compute) if compute is not None else int(keep)
        self.real = real
        self.ports = list(ports)

    def modes(self):
        out = []
        for c in self.children:
            out += [c[0]] if isinstance(c, tuple) else c.modes()
        return out

    def child_of(self, mode):
        for a, c in enumerate(self.children):
            if (isinstance(c, tuple) and c[0] == mode) or (isinstance(c, Node) and mode in c.modes()):
                return a
        raise KeyError(mode)

    def leaf_basis(self, mode):
        c = self.children[self.child_of(mode)]
        return c[1] if isinstance(c, tuple) else c.leaf_basis(mode)


@dataclass
class MaterializedNode:
    node: Node
    dims: list
    energies: np.ndarray
    vectors: np.ndarray
    children: list
    operator: KronFreeOperator = None
    solver: object = None
    stats: dict = field(default_factory=dict)
    seconds: float = 0.0
    ports: dict = field(default_factory=dict)   # port matrices of basis children, by key


def _axis_apply(op, V, dims, a):
    """op applied along axis a of the column-major tensors V[:, j] (host)."""
    if len(dims) == 1:
        return dense(op) @ V
    T = V.reshape(list(dims) + [V.shape[1]], order="F")
    T = np.moveaxis(np.tensordot(dense(op), np.moveaxis(T, a, 0), axes=(1, 0)), 0, a)
    return T.reshape(V.shape, order="F")


# Port formation moves to the device (kfc_sandwich) from this many entries of V on
# (the zero-pi parent's 48 000 x 16 block is above it, the leaves are below).
DEVICE_PORT_ENTRIES = 1 << 19


def sandwich(V, dims, ops, ctx=None):
    """V' (x_a O_a) V on the device (``kfc_sandwich``): ``ops`` is a list of
    (axis, LocalOp); the operator is built as one AxisTerm and applied to V
    uploaded once, then contracted with V on the matrix cores.  A complex
    operator on real V is taken in complex arithmetic (the reference forms its
    real and imaginary components separately, src/materialize.jl:456-461; the
    result is the same matrix)."""
    import ctypes as C
    import torch
    V = np.asarray(V)
    N, k = V.shape
    cplx = np.iscomplexobj(V) or any(np.iscomplexobj(dense(op)) for _, op in ops)
    term = AxisTerm(1.0 + 0j if cplx else 1.0, sorted(ops, key=lambda t: t[0]))
    H = KronFreeOperator(list(dims), [term], constant=0.0, ctx=ctx)
    if (H.dtype == np.complex128) != cplx:
        raise L.ArgumentError("sandwich: operator and block types disagree")
    Vt = torch.from_numpy(np.ascontiguousarray(V.T.astype(np.complex128 if cplx else np.float64))).to("cuda")
    G = np.empty(2 * k * k if cplx else k * k)
    H.ctx.wait_torch()
    L.check(L.lib().kfc_sandwich(H.handle, C.c_void_p(Vt.data_ptr()), int(k), int(N), G.ctypes.data_as(L._D)),
            H.ctx.handle)
    if cplx:
        G = G.view(np.complex128)
    return G.reshape((k, k), order="F")


def _basis_key(b):
    """A basis's type and scalar parameters (None when it holds anything else)."""
    items = []
    for k, v in sorted(vars(b).items()):
        if not isinstance(v, (bool, int, float, str)):
            return None
        items.append((k, v))
    return (type(b).__name__,) + tuple(items)


def port_matrix(mn, mode, port, device=None):
    """``_port_matrix`` (src/materialize.jl:443-459): V' O V with O the mode's port
    on its axis (a basis child's localop or a node child's port).  ``device``:
    form it with ``sandwich`` (default: when V has at least DEVICE_PORT_ENTRIES
    entries).  Ports on basis children are kept on the materialized node (its
    vectors never change), keyed by mode, port and the basis's parameters, so a node
    shared by many parents (a sweep's fixed leaf) forms each port once."""
    a = mn.node.child_of(mode)
    ch = mn.node.children[a]
    key = None
    if isinstance(ch, tuple):
        bk = _basis_key(ch[1])
        if bk is not None:
            key = (mode, port, bk, device)
            try:
                hash(key)
            except TypeError:
                key = None
    if key is not None and key in mn.ports:
        return mn.ports[key].copy()
    op = localop(ch[1], port) if isinstance(ch, tuple) else DenseOp(port_matrix(mn.children[a], mode, port))
    V = mn.vectors
    if device is None:
        device = V.size >= DEVICE_PORT_ENTRIES
    if device:
        P = sandwich(V, mn.dims, [(a, op)])
    else:
        P = V.conj().T @ _axis_apply(op, V, mn.dims, a)
    if key is not None:
        mn.ports[key] = P.copy()
    return P


def joint_port(mn, factors, device=None):
    """``_joint_port`` (src/materialize.jl:490-518): V' (x_a O_a) V."""
    V = mn.vectors
    ops = []
    for mode, port in factors:
        a = mn.node.child_of(mode)
        ch = mn.node.children[a]
        op = localop(ch[1], port) if isinstance(ch, tuple) else DenseOp(port_matrix(mn.children[a], mode, port))
        ops.append((a, op))
    if device is None:
        device = V.size >= DEVICE_PORT_ENTRIES
    if device:
        return sandwich(V, mn.dims, ops)
    W = V
    for a, op in ops:
        W = _axis_apply(op, W, mn.dims, a)
    return V.conj().T @ W


def _hermitian(M):
    M = np.asarray(M)
    return np.allclose(M, M.conj().T, rtol=1e-10, atol=1e-12 * max(1.0, float(np.max(np.sum(np.abs(M), axis=1)))))


def port_operator(mn, key, ctx=None):
    """``port_operator(sys, key, node)`` (src/portoperator.jl:22-39): the port
    ``key = (mode, port)``, or the joint port ``[(mode, port), ...]`` of factors on
    distinct modes, as a KronFreeOperator on the product basis of the materialized
    node ``mn`` (its children's kept states, axis 0 fastest), identity on the axes the
    key does not touch.  Each touched axis carries its child's operator: a basis
    child's localop, or a node child's port (its joint port when several factors fall
    in that child).  Real exactly when every factor is real; Hermitian when every
    factor is, to rounding."""
    keys = [key] if isinstance(key, tuple) else list(key)
    if not keys:
        raise L.ArgumentError("port_operator: the key is empty")
    modes = mn.node.modes()
    for m, _ in keys:
        if m not in modes:
            raise L.ArgumentError(f"mode :{m} is not a mode of the node on modes {modes}")
    if len({m for m, _ in keys}) != len(keys):
        raise L.ArgumentError(f"port_operator: the factors of a joint key must be on distinct modes, got {keys}")
    groups = {}
    for m, p in keys:
        groups.setdefault(mn.node.child_of(m), []).append((m, p))
    ops = []
    for a in sorted(groups):
        ch, fs = mn.node.children[a], groups[a]
        if isinstance(ch, tuple):
            op = localop(ch[1], fs[0][1])
        elif len(fs) == 1:
            op = DenseOp(port_matrix(mn.children[a], *fs[0]))
        else:
            op = DenseOp(joint_port(mn.children[a], fs))
        ops.append((a, op))
    cplx = any(np.iscomplexobj(dense(op)) for _, op in ops)
    herm = all(_hermitian(dense(op)) for _, op in ops)
    term = AxisTerm(1.0 + 0j if cplx else 1.0, ops)
    return KronFreeOperator(list(mn.dims), [term], constant=0.0, hermitian=herm, ctx=ctx)


def node_axisterms(node, kids, terms, real):
    """The AxisTerms of a node (``_node_operator`` / ``_real_node_terms``,
    src/materialize.jl:583-647): on a real node every routed term's
    Re(c (x) O) expansion with the real terms on equal (axis, canonical
    sub-key, component) keys summed; each node child's kept energies as a
    DiagonalOp on its axis."""
    out = []
    index, sums, firsts = {}, [], []
    for c, fs in terms:
        groups = {}
        for mode, port in fs:
            groups.setdefault(node.child_of(mode), []).append((mode, port))
        ops = []
        for a in sorted(groups):
            g = groups[a]
            ch = node.children[a]
            if len(g) == 1:
                mode, port = g[0]
                op = localop(ch[1], port) if isinstance(ch, tuple) else DenseOp(port_matrix(kids[a], mode, port))
                kind = character(node.leaf_basis(mode), port)
                cport, sgn = canonical_port(port, kind)
                ops.append((a, op, kind, sgn, (("single", mode, cport),)))
            else:
                op = DenseOp(joint_port(kids[a], g))
                kinds = [character(node.leaf_basis(mm), pp) for mm, pp in g]
                kind = "paired" if "paired" in kinds else ("imaginary" if kinds.count("imaginary") % 2 else "real")
                ops.append((a, op, kind, 1, (("joint",) + tuple(sorted((mm, pp) for mm, pp in g)),)))
        if not real:
            out.append(AxisTerm(complex(c), [(a, op) for a, op, _, _, _ in ops]))
            continue
        comps = []
        for a, op, kind, sgn, key in ops:
            R = None if kind == "imaginary" else realpart(op)
            I = None if kind == "real" else imagpart(op)
            ch = (False,) if kind == "real" else ((True,) if kind == "imaginary" else (False, True))
            comps.append((a, sgn, R, I, ch, key))
        for S in itertools.product(*[x[4] for x in comps]):
            coeff = _re_ipow(c, sum(S))
            if coeff == 0:
                continue
            sign = 1
            for x, inS in zip(comps, S):
                if inS:
                    sign *= x[1]
            key = tuple((x[0], x[5], inS) for x, inS in zip(comps, S))
            facs = [(x[0], x[3] if inS else x[2]) for x, inS in zip(comps, S)]
            if key in index:
                sums[index[key]] += coeff * sign
            else:
                index[key] = len(sums)
                sums.append(coeff * sign)
                firsts.append((sign, facs))
    for s, (sign, facs) in zip(sums, firsts):
        if s != 0:
            out.append(AxisTerm(s * sign, facs))
    for a, ch in enumerate(node.children):
        if isinstance(ch, Node):
            out.append(AxisTerm(1.0, [(a, DiagonalOp(np.asarray(kids[a].energies, float)))]))
    return out


def route(root, terms):
    """``route`` (src/tree.jl:500-529): each term to the lowest node containing its modes."""
    routed = {}

    def lowest(nd, modes):
        for ch in nd.children:
            if isinstance(ch, Node) and all(m in ch.modes() for m in modes):
                return lowest(ch, modes)
        return nd
    for t in terms:
        routed.setdefault(id(lowest(root, [f[0] for f in t[1]])), []).append(t)
    return routed


def positive_mask(node, dims):
    """The phase mask of a node's product basis (src/materialize.jl:470-472,
    src/bases.jl positive_mask): a basis child's coordinates >= 0, all of a node
    child's kept states; axis 0 fastest."""
    mask = np.ones(1, dtype=bool)
    for ch, d in zip(node.children, dims):
        if isinstance(ch, tuple):
            b = ch[1]
            if hasattr(b, "coordinates"):
                m = b.coordinates() >= 0
            elif hasattr(b, "ncut"):
                m = np.arange(-b.ncut, b.ncut + 1) >= 0
            else:
                m = np.ones(d, dtype=bool)
        else:
            m = np.ones(d, dtype=bool)
        mask = np.kron(m, mask).astype(bool)
    return mask


def fix_phases(V, mask):
    """``fix_phases!(V, :positive_sector, mask)`` (src/phase.jl:12-33): the
    largest-modulus entry among the mask made real and positive (the global
    largest when the column has no weight there)."""
    V = np.array(V)
    sector = np.flatnonzero(mask)
    for j in range(V.shape[1]):
        col = V[:, j]
        i = sector[np.argmax(np.abs(col[sector]))] if len(sector) else 0
        if len(sector) == 0 or abs(col[i]) < 1e-14:
            i = int(np.argmax(np.abs(col)))
        x = col[i]
        if np.iscomplexobj(V):
            V[:, j] = col * np.exp(-1j * np.angle(x))
        elif np.real(x) < 0:
            V[:, j] = -col
    return V


class System:
    """A materialized tree: ``energies(sys)``, ``port(sys, mode, port)``."""

    def __init__(self, root):
        self.root = root

    def energies(self):
        return self.root.energies

    def port(self, mode, port):
        return port_matrix(self.root, mode, port)


def operator(root, spec, kids=None, ctx=None, block_capacity=1):
    """The root node's KronFreeOperator with its children materialized, not
    solved (``operator(sys)``, src/materialize.jl:982-992).  ``kids`` may pass
    already materialized children (a sweep reuses the ones that do not depend
    on the swept parameters); ``ctx`` picks the device context."""
    routed = route(root, circuit_terms(spec))
    if kids is None:
        kids = []
        for ch in root.children:
            if isinstance(ch, tuple):
                kids.append(None)
                continue
            sub = materialize(ch, spec, _routed=routed, _is_root=False)
            kids.append(sub.root)
    dims = [ch[1].dim if isinstance(ch, tuple) else len(k.energies) for ch, k in zip(root.children, kids)]
    ts = node_axisterms(root, kids, routed.get(id(root), []), root.real)
    return KronFreeOperator(dims, ts, constant=constant_term(spec), ctx=ctx, block_capacity=block_capacity), kids


def materialize_child(root, child, spec):
    """One child node of ``root`` materialized with the terms ``route`` gives it."""
    routed = route(root, circuit_terms(spec))
    return materialize(child, spec, _routed=routed, _is_root=False).root


def materialize(root, spec, previous=None, X0=None, _routed=None, _is_root=True):
    """``materialize(root, spec, theta; previous)`` (src/materialize.jl:177-192, 713-731):
    children first, then each node's operator and solve on the device.
    ``X0`` (a dict node -> block) supplies LOBPCG warm starts directly."""
    import time
    routed = _routed if _routed is not None else route(root, circuit_terms(spec))
    const = constant_term(spec)
    prev_map = {}
    if previous is not None:
        def walk(a, b):
            prev_map[id(a)] = b
            for ca, cb in zip(a.children, b.children):
                if isinstance(ca, Node):
                    walk(ca, cb)
        walk(root, previous.root)

    def visit(nd, is_root):
        kids = [None if isinstance(ch, tuple) else visit(ch, False) for ch in nd.children]
        dims = [ch[1].dim if isinstance(ch, tuple) else len(k.energies) for ch, k in zip(nd.children, kids)]
        t0 = time.perf_counter()
        m = min(nd.compute, int(np.prod(dims)))
        s = resolve(nd.solver, {"isleaf": all(isinstance(c, tuple) for c in nd.children), "dims": dims, "m": m})
        ts = node_axisterms(nd, kids, routed.get(id(nd), []), nd.real)
        cap = m + s.guard if isinstance(s, LOBPCG) else 1
        H = KronFreeOperator(dims, ts, constant=const if is_root else 0.0, block_capacity=max(1, cap))
        x0 = None
        if isinstance(s, LOBPCG):
            if X0 is not None and id(nd) in X0:
                x0 = X0[id(nd)]
            elif id(nd) in prev_map and list(prev_map[id(nd)].dims) == dims:
                pv = prev_map[id(nd)].vectors           # _initial_block (src/materialize.jl:656-666)
                k = min(pv.shape[1], m)
                x0 = np.concatenate([pv[:, :k], np.random.default_rng(0).standard_normal((pv.shape[0], m - k))], 1)
        E, V, st = solve_stats(s, H, m if not isinstance(s, Dense) else m, x0)
        keep = min(nd.keep, len(E))
        V = fix_phases(np.asarray(V)[:, :keep], positive_mask(nd, dims))
        return MaterializedNode(nd, dims, np.asarray(E)[:keep], V, kids, H, s, st, time.perf_counter() - t0)
    return System(visit(root, _is_root))
