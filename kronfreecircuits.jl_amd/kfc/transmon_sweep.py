"""Batched sweep of coupled-transmon instances on periodic phase grids.

An instance is the two-transmon circuit (test/fixtures/two_transmons.json form:
E_C matrix [[E_C1, g/8], [g/8, E_C2]], cosines E_J1, E_J2, offset charges n_g1,
n_g2) with each mode on PeriodicGrid(N_i, order), N_i drawn from {256, 384, 512,
640, 768}: two complex leaves (the offset charge makes n complex) keeping 8
states, and a top node [8, 8] keeping 20.  There is no reference counterpart
for the batch (the closest is the warm-started sweep of bench/common.jl:135-163);
each instance is the tree that materialize (src/materialize.jl:713-731) would
solve, and its energies are the ones it returns.

Two ways to run a batch:

* plain: instance after instance, every leaf by FP64 Lanczos, the top Dense
  (what materialize does per point on the device);
* batched: the leaves of all instances in flight together, one host thread per
  device context, then the tops.

The leaves are solved by Lanczos: a periodic grid of N points with an order-8
stencil has a kinetic spectrum reaching 4 E_C (N/2)^2 ~ 1e5-1e6 GHz, where
LOBPCG with a diagonal preconditioner (nearly constant here) needs thousands of
iterations while a Krylov method needs a few hundred matvecs.  The leaf
tolerance is 1e-9, above the rounding floor 10 eps |H| of these spectra, and
the Krylov dimension 48 (the reference's krylovdim keyword; its default
2m + 10 = 26 restarts four times as often on these spectra)."""
from __future__ import annotations

import time

import numpy as np

from . import _lib as L
from .bases import PeriodicGrid
from .operator import KronFreeOperator
from concurrent.futures import ThreadPoolExecutor

from .solvers import ChebyshevFilter, Dense, Lanczos, solve_chebyshev_batch, solve_dense_batch, solve_stats
from .tree import (MaterializedNode, Node, circuit_terms, constant_term, fix_phases, materialize, node_axisterms,
                   operator, positive_mask, route)
from .configs import two_transmons_spec

GRID_SIZES = (256, 384, 512, 640, 768)
# the ranges the instance parameters are drawn from (GHz): about the fixture's
# values (E_J 25-30, E_C 1.0-1.2, g 0.3)
RANGES = dict(E_J=(20.0, 35.0), E_C=(0.8, 1.4), n_g=(0.0, 1.0), g=(0.1, 0.5))


def draw(n, seed=20261203):
    """n instances: grid sizes and parameters, seeded."""
    r = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        u = lambda k: float(r.uniform(*RANGES[k]))
        out.append(dict(N1=int(r.choice(GRID_SIZES)), N2=int(r.choice(GRID_SIZES)), E_J1=u("E_J"), E_C1=u("E_C"),
                        n_g1=u("n_g"), E_J2=u("E_J"), E_C2=u("E_C"), n_g2=u("n_g"), g=u("g")))
    return out


def spec_of(inst):
    return two_transmons_spec(inst["E_J1"], inst["E_C1"], inst["n_g1"], inst["E_J2"], inst["E_C2"], inst["n_g2"],
                              inst["g"])


def tree_of(inst, order=8, k_leaf=8, k_top=20, leaf_solver=None):
    # the offset charges make every node complex (n on a periodic grid is imaginary)
    leaves = [Node([(f"varphi_{i}", PeriodicGrid(inst[f"N{i}"], order))], keep=k_leaf, solver=leaf_solver,
                   real=False) for i in (1, 2)]
    return Node(leaves, keep=k_top, solver=Dense(), real=False)


class TransmonSweep:
    def __init__(self, order=8, k_leaf=8, k_top=20, contexts=8):
        self.order, self.k_leaf, self.k_top = order, k_leaf, k_top
        self.ctxs = [L.context(0)] + [L.Context(0) for _ in range(max(0, contexts - 1))]

    def run_plain(self, insts, tol=1e-9, mode="f64"):
        """Instance after instance: materialize with Lanczos leaves and a Dense top."""
        out, t0 = [], time.perf_counter()
        for inst in insts:
            root = tree_of(inst, self.order, self.k_leaf, self.k_top, Lanczos(tol=tol, krylovdim=48, mode=mode))
            out.append(materialize(root, spec_of(inst)).root)
        return out, time.perf_counter() - t0

    def run_batched(self, insts, mode="f64", tol=1e-9, method="lanczos", degree=60, extra=8):
        """All leaves of all instances solved together, then every top.  method
        "lanczos": one host thread per device context, each working through its
        share of the leaves; "chebyshev": one batched Chebyshev-filtered subspace
        iteration over every leaf on one context (kfc_chebyshev_batch), each
        filter step one launch for the whole batch."""
        t0 = time.perf_counter()
        roots, specs, leaf_ops, leaf_nodes = [], [], [], []
        c = 0
        for inst in insts:
            root = tree_of(inst, self.order, self.k_leaf, self.k_top)
            spec = spec_of(inst)
            rt = route(root, circuit_terms(spec))
            roots.append(root)
            specs.append(spec)
            for leaf in root.children:
                dims = [leaf.children[0][1].dim]
                ts = node_axisterms(leaf, [None], rt.get(id(leaf), []), leaf.real)
                ctx = self.ctxs[0] if method == "chebyshev" else self.ctxs[c % len(self.ctxs)]
                H = KronFreeOperator(dims, ts, constant=0.0, ctx=ctx)
                c += 1
                leaf_ops.append(H)
                leaf_nodes.append(leaf)
        nctx = len(self.ctxs)
        if method == "chebyshev":
            solver = ChebyshevFilter(tol=tol, degree=degree, extra=extra, mode=mode)
            res, st = solve_chebyshev_batch(leaf_ops, self.k_leaf, extra=extra, degree=degree, tol=tol, mode=mode)
            results = [(E, V, s_) for (E, V), s_ in zip(res, st)]
        else:
            solver = Lanczos(tol=tol, krylovdim=48, mode=mode)

            def work(k):   # every nctx-th leaf: the ones bound to context k
                return [(i, solve_stats(solver, leaf_ops[i], self.k_leaf)) for i in range(k, len(leaf_ops), nctx)]
            results = [None] * len(leaf_ops)
            with ThreadPoolExecutor(max_workers=nctx) as ex:
                for part in ex.map(work, range(nctx)):
                    for i, r in part:
                        results[i] = r
        t_leaves = time.perf_counter() - t0
        mats = []
        for (E, V, st), H, leaf in zip(results, leaf_ops, leaf_nodes):
            V = fix_phases(np.asarray(V), positive_mask(leaf, H.dims))
            mats.append(MaterializedNode(leaf, list(H.dims), np.asarray(E), V, [None], H, solver, st))
        # every top's operator, then their Dense solves at once (host threads)
        Hts = [operator(root, spec, kids=mats[2 * i:2 * i + 2])[0] for i, (root, spec) in enumerate(zip(roots, specs))]
        res, sts = solve_dense_batch(Hts, self.k_top)
        tops = [MaterializedNode(root, list(Ht.dims), np.asarray(E), np.asarray(V), mats[2 * i:2 * i + 2], Ht, Dense(), st)
                for i, (root, Ht, (E, V), st) in enumerate(zip(roots, Hts, res, sts))]
        return tops, time.perf_counter() - t0, t_leaves
