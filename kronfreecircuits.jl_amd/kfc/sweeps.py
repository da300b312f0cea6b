"""Parameter sweeps of the zero-pi tree and its learned warm start.

A sweep point is one draw of (E_CS, n_g) of the zero-pi tree (the workload
described in kfc.configs); its parent node [40, 40, 30] is solved by LOBPCG for
the lowest 16 pairs.  Three ways to run a sweep:

* plain: one point after another, FP64 throughout, product-state seed (the
  reference's materialize per point);
* batched: the parent solves of many points in flight at once on several
  device contexts (kfc_lobpcg_batch), FP32 bulk iterations with FP64
  Rayleigh-Ritz and refinement (mode="mixed");
* learned: batched, plus an initial block X0 predicted from the parameters by
  the warm-start network;
* learned subspace: batched, each point started from the lowest Ritz vectors of
  its operator on a learned subspace (LearnedSubspace).

Both learned starts work in a canonical frame of the parent basis that does not
move with the parameters: the theta leaf's kept states are charge states
|n>, whose energy order changes with n_g, so a parent vector is re-indexed by
(charge n, phi level, zeta level).  The window holds every charge the theta
leaf keeps at any n_g in [0, 1] and every phi and zeta level, so the lowest
parent states lie inside it.  A fixed basis B of that frame is taken from the
training eigenvectors (their leading principal directions).  The network maps
the six circuit parameters to the coefficients of the lowest m + g parent
eigenvectors on B (X0 = B C mapped back to the parent basis of the point); the
learned subspace instead hands span(B) itself to a Rayleigh-Ritz step on the
device, whose lowest m + g Ritz vectors start LOBPCG."""
from __future__ import annotations

import math
import time

import numpy as np

from . import _lib as L
from . import configs
from .solvers import LOBPCG, solve_batch, solve_batch_subspace, solve_stats
from .tree import materialize_child, operator

WINDOW = (-19, 20)       # charges n in [-19, 20): kept by the theta leaf (40 states) at every n_g in [0, 1]
PHI_LEVELS = 40
ZETA_LEVELS = 30


def principal_directions(Z, k, seed=0):
    """The leading k left singular vectors of Z (frame rows x training vectors), an
    orthonormal frame basis (randomized range finder with power iterations on the
    device when one is present; the spectra here fall off by orders of magnitude)."""
    import torch
    dev = "cuda" if torch.cuda.is_available() else "cpu"
    Zt = torch.as_tensor(np.ascontiguousarray(Z), dtype=torch.float64, device=dev)
    k = min(k, min(Zt.shape))
    g = torch.Generator(device=dev).manual_seed(seed)
    Q = Zt @ torch.randn(Zt.shape[1], min(Zt.shape[1], k + 32), dtype=torch.float64, device=dev, generator=g)
    for _ in range(4):
        Q, _ = torch.linalg.qr(Q)
        Q, _ = torch.linalg.qr(Zt @ (Zt.T @ Q))
    U, S, _ = torch.linalg.svd(Q.T @ Zt, full_matrices=False)
    return (Q @ U[:, :k]).cpu().numpy(), S[:k].cpu().numpy()


def features(theta):
    """The six network inputs: (E_CS, n_g, E_J, E_L, dE_J, varphi_ext / 2pi), each
    mapped to about [-1, 1] over its range (E_J, E_L, dE_J and varphi_ext are
    held at their fixed values in these sweeps)."""
    return np.array([(theta["E_CS"] - 0.11) / 0.09, (theta["n_g"] - 0.5) / 0.5, (theta.get("E_J", 10.0) - 10.0),
                     (theta.get("E_L", 0.05) - 0.05) / 0.01, (theta.get("dE_J", 0.05) - 0.05) / 0.02,
                     (theta.get("varphi_ext", 2 * math.pi * 0.23) / (2 * math.pi) - 0.23) / 0.05], dtype=np.float32)


class ZeroPiSweep:
    def __init__(self, ncut=40, Nphi=256, Nzeta=30, keep=40, m=16, contexts=4, guard=4):
        self.ncut, self.Nphi, self.Nzeta, self.keep, self.m = ncut, Nphi, Nzeta, keep, m
        self.guard = guard
        self.ctxs = [L.context(0)] + [L.Context(0) for _ in range(max(0, contexts - 1))]
        self._phi = None

    def tree(self, theta):
        spec, E_Cz = configs.zero_pi_spec(theta["E_CS"], theta["n_g"])
        root = configs.zero_pi_tree(E_Cz, ncut=self.ncut, Nphi=self.Nphi, Nzeta=self.Nzeta, keep=self.keep, m=self.m)
        return spec, root

    def parent(self, theta, ctx=None):
        """The parent operator of one sweep point (the phi leaf, which no swept
        parameter touches, is materialized once)."""
        spec, root = self.tree(theta)
        th = materialize_child(root, root.children[0], spec)
        if self._phi is None:
            self._phi = materialize_child(root, root.children[1], spec)
        ph = self._phi
        ph.node = root.children[1]
        ze = materialize_child(root, root.children[2], spec)
        H, kids = operator(root, spec, kids=[th, ph, ze], ctx=ctx, block_capacity=self.m + 4)
        return H, kids

    # -- the canonical frame ---------------------------------------------------------------
    def frame_index(self, kids):
        """Parent-basis flat index of every canonical (n, phi level, zeta level) in the
        window, or -1 when that charge is not among the kept theta states."""
        Vt = kids[0].vectors
        charge = np.argmax(np.abs(Vt), axis=0) - self.ncut          # kept index -> charge n
        d0, d1, d2 = Vt.shape[1], kids[1].vectors.shape[1], kids[2].vectors.shape[1]
        lo, hi = WINDOW
        kept = np.full(hi - lo, -1, dtype=np.int64)                 # charge n -> kept index
        inside = (charge >= lo) & (charge < hi)
        kept[charge[inside] - lo] = np.flatnonzero(inside)
        nz, nphi = min(ZETA_LEVELS, d2), min(PHI_LEVELS, d1)
        # frame order: charge fastest, then phi level, then zeta level
        k = np.broadcast_to(kept[None, None, :], (nz, nphi, hi - lo))
        plane = d0 * (np.arange(nphi)[None, :, None] + d1 * np.arange(nz)[:, None, None])
        return np.where(k >= 0, k + plane, -1).reshape(-1)

    def to_frame(self, V, idx):
        out = np.zeros((len(idx), V.shape[1]))
        ok = idx >= 0
        out[ok] = V[idx[ok]]
        return out

    def from_frame(self, Xf, idx, N):
        X = np.zeros((N, Xf.shape[1]))
        ok = idx >= 0
        X[idx[ok]] = Xf[ok]
        return X

    # -- sweeps ----------------------------------------------------------------------------
    def run_plain(self, thetas, tol=1e-8):
        """One point after another, FP64, product seed."""
        out, t0 = [], time.perf_counter()
        for th in thetas:
            H, kids = self.parent(th)
            E, V, st = solve_stats(LOBPCG(tol=tol, mode="f64"), H, self.m)
            out.append((E, V, st, kids))
        return out, time.perf_counter() - t0

    def run_plain_warm(self, thetas, tol=1e-8):
        """One point after another, FP64, each warm-started from the previous point's
        kept vectors as the reference's sweeps do (``materialize(...; previous)`` ->
        ``_initial_block``, src/materialize.jl:656-666, bench/common.jl:135-163): the
        raw coefficients of the previous parent solution, m columns, the guards from
        the product seed; the first point starts cold."""
        out, t0, prev = [], time.perf_counter(), None
        for th in thetas:
            H, kids = self.parent(th)
            X0 = prev if prev is not None and prev.shape[0] == H.N else None
            E, V, st = solve_stats(LOBPCG(tol=tol, mode="f64"), H, self.m, X0=X0)
            out.append((E, V, st, kids))
            prev = V[:, :self.m]
        return out, time.perf_counter() - t0

    def width(self, N):
        """Columns LOBPCG converges on a parent of N states: m plus the guards."""
        return self.m + max(0, min(self.guard, (N - 5) // 3 - self.m))

    def run_batched(self, thetas, mode="mixed", tol=1e-8, x0_fn=None, nev=None, lockstep=False, subspace=None,
                    guard=None):
        """All points' parent solves in flight: on the contexts round-robin, or
        (lockstep) on one context advancing together; ``nev`` pairs per point
        (default m), ``guard`` guard columns (default the sweep's).  Lockstep serves
        blocks of up to 21 columns (m + guards); a wider request runs on the
        contexts.  ``subspace`` (a fitted LearnedSubspace): lockstep, every point
        started from the lowest Ritz vectors on the learned subspace."""
        n_ev = nev or self.m
        gd = self.guard if guard is None else guard
        if subspace is not None:
            lockstep = True
        if lockstep and 3 * (n_ev + gd) > 64:
            if subspace is not None:
                raise L.UnsupportedError("a learned-subspace sweep runs in lockstep: m + guard must stay below 22")
            lockstep = False
        t0 = time.perf_counter()
        Hs, kidss = [], []
        for i, th in enumerate(thetas):
            H, kids = self.parent(th, ctx=self.ctxs[0 if lockstep else i % len(self.ctxs)])
            Hs.append(H)
            kidss.append(kids)
        X0s = None
        if x0_fn is not None:
            X0s = x0_fn(thetas, Hs, kidss)
        solver = LOBPCG(tol=tol, mode=mode, guard=gd)
        if subspace is not None:
            res, st = solve_batch_subspace(solver, Hs, n_ev, subspace.B, [self.frame_index(k) for k in kidss])
        else:
            res, st = solve_batch(solver, Hs, n_ev, X0s, lockstep=lockstep)
        wall = time.perf_counter() - t0
        return [(E, V, s, k) for (E, V), s, k in zip(res, st, kidss)], wall


class WarmStart:
    """The learned warm start: a fixed 64-vector basis of the canonical frame and
    the network mapping parameters to the coefficients on it of the lowest m + g
    parent eigenvectors (LOBPCG converges the g guard columns too, so a start
    that leaves them to the product seed gains nothing: 64 x 20 outputs for the
    zero-pi parent's 16 + 4)."""

    def __init__(self, sweep, nbasis=64, hidden=256, seed=0, width=None):
        from .mlp import MLP
        self.sw, self.nb = sweep, nbasis
        self.m = width or sweep.m + sweep.guard
        self.net = MLP((6, hidden, hidden, nbasis * self.m), seed=seed)
        self.B = None
        self.signref = None

    def fit_basis(self, frames):
        """B: the leading principal directions of the stacked training eigenvectors."""
        self.B, _ = principal_directions(np.concatenate([f[:, :self.m] for f in frames], axis=1), self.nb)

    def coeffs(self, Vf):
        C = self.B.T @ Vf[:, :self.m]           # nb x m
        if self.signref is None:
            self.signref = np.argmax(np.abs(C), axis=0)
        s = np.sign(C[self.signref, np.arange(C.shape[1])])
        s[s == 0] = 1
        return C * s

    def train(self, thetas, Vframes, steps=3000, lr=1e-3, batch=256):
        self.fit_basis(Vframes)
        X = np.stack([features(t) for t in thetas])
        T = np.stack([self.coeffs(Vf).reshape(-1, order="F") for Vf in Vframes]).astype(np.float32)
        self.scale = float(np.max(np.abs(T))) or 1.0
        loss0 = float(np.mean((self.net.forward(X) - T / self.scale) ** 2))
        loss = self.net.train(X, T / self.scale, batch=min(batch, len(X)), steps=steps, lr=lr)
        return loss0, loss

    def predict_frames(self, thetas):
        X = np.stack([features(t) for t in thetas])
        Y = self.net.forward(X) * self.scale
        return [self.B @ y.reshape(self.nb, self.m, order="F") for y in Y]

    def x0_fn(self, thetas, Hs, kidss):
        out = []
        for Xf, H, kids in zip(self.predict_frames(thetas), Hs, kidss):
            X0 = self.sw.from_frame(Xf, self.sw.frame_index(kids), H.N)[:, :self.sw.width(H.N)]
            bad = np.linalg.norm(X0, axis=0) == 0
            if np.any(bad):
                X0[:, bad] = np.random.default_rng(0).standard_normal((H.N, int(bad.sum())))
            out.append(X0)
        return out


class LearnedSubspace:
    """The learned subspace: the leading ``nbasis`` principal directions of the
    training eigenvectors in the canonical frame (kept on the device).  A sweep
    started from it (``run_batched(..., subspace=...)``, kfc_lobpcg_lockstep_subspace)
    begins each point at the lowest m + g Ritz vectors of that point's operator on
    span(B): the Rayleigh-Ritz step resolves the order and mixing of neighbouring
    states that a regression of individual eigenvectors blurs, so the start is
    accurate in exactly the directions (the next states up) that set LOBPCG's
    convergence of the last block columns."""

    def __init__(self, sweep, nbasis=256):
        self.sw, self.nb = sweep, nbasis
        self.B = None
        self.singular_values = None

    def fit(self, frames, ncols=None):
        """B from the training eigenvectors in the frame (their first ``ncols``
        columns each, default all)."""
        import torch
        Z = np.concatenate([f if ncols is None else f[:, :ncols] for f in frames], axis=1)
        U, S = principal_directions(Z, self.nb)
        self.B = torch.tensor(U, dtype=torch.float64, device="cuda")
        self.singular_values = S
        return self


def path_order(thetas):
    """A walk through the points, each next point the nearest unvisited one in
    (E_CS, n_g) scaled to their ranges, from the smallest E_CS: the order that gives a
    sweep warm-started from the previous point its closest neighbours."""
    P = np.array([[(t["E_CS"] - 0.02) / 0.18, t["n_g"]] for t in thetas])
    left = list(range(len(thetas)))
    cur = int(np.argmin(P[:, 0]))
    order = [cur]
    left.remove(cur)
    while left:
        d = np.linalg.norm(P[left] - P[cur], axis=1)
        cur = left[int(np.argmin(d))]
        order.append(cur)
        left.remove(cur)
    return order


def subspace_capture(X0, V):
    """The weight of each true eigenvector in span(X0) (1 = captured)."""
    Q, _ = np.linalg.qr(X0)
    return np.sum((Q.T @ V) ** 2, axis=0)
