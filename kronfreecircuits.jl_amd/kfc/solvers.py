"""The Solver plugin surface (src/solvers.jl:3-31) and its keyword structs
(src/tree.jl:6-351), served by libkfc on the device.

``solve(s, H, m; X0)`` returns ``(E, V)`` on the host with ``E`` ascending and
``V`` of shape ``(N, m)``; ``solve_stats`` also returns the work counters.
Every solver accepts ``mode="f64"`` (the plain FP64 path) or ``mode="mixed"``
(FP32 applies for the bulk iteration, FP64 Rayleigh-Ritz, reorthogonalization
and a final FP64 refinement), a solver field added beside the reference's
unchanged keywords."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .operator import pack_block, unpack_block

_MODES = {"f64": L.KFC_MODE_F64, "mixed": L.KFC_MODE_MIXED}


def _mode(m):
    if m not in _MODES:
        raise L.ArgumentError(f"mode must be 'f64' or 'mixed', got {m!r}")
    return _MODES[m]


class Solver:
    pass


class Dense(Solver):
    """Full dense diagonalization (src/tree.jl:4)."""

    def __repr__(self):
        return "Dense()"


class Lanczos(Solver):
    """``Lanczos(; tol=1e-10, krylovdim=0, maxiter=1000, verbosity=0, on)`` (src/tree.jl:6-23)."""

    def __init__(self, tol=1e-10, krylovdim=0, maxiter=1000, verbosity=0, on=None, mode="f64", seed=0x4c5a):
        self.tol, self.krylovdim, self.maxiter, self.verbosity, self.on = tol, krylovdim, maxiter, verbosity, on
        self.mode, self.seed = mode, seed
        _mode(mode)


class LOBPCG(Solver):
    """``LOBPCG(; tol=1e-8, maxiter=1000, preconditioner=:ritz, eta=0.01, guard=4,
    seed=:product, on, callback)`` (src/tree.jl:66-87)."""

    def __init__(self, tol=1e-8, maxiter=1000, preconditioner="ritz", eta=0.01, guard=4, seed="product", on=None,
                 callback=None, mode="f64", rng_seed=0x4b46):
        if not (tol > 0 and math.isfinite(tol)):
            raise L.ArgumentError(f"LOBPCG: tol must be positive and finite, got {tol}")
        if maxiter < 1:
            raise L.ArgumentError(f"LOBPCG: maxiter must be at least 1, got {maxiter}")
        if preconditioner not in ("ritz", "jacobi", "none"):
            raise L.ArgumentError(f"LOBPCG: preconditioner must be :ritz, :jacobi or :none, got :{preconditioner}")
        if not (eta > 0 and math.isfinite(eta)):
            raise L.ArgumentError(f"LOBPCG: eta must be positive and finite, got {eta}")
        if guard < 0:
            raise L.ArgumentError(f"LOBPCG: guard must be nonnegative, got {guard}")
        if seed not in ("product", "random"):
            raise L.ArgumentError(f"LOBPCG: seed must be :product or :random, got {seed!r}")
        self.tol, self.maxiter, self.preconditioner, self.eta = tol, maxiter, preconditioner, eta
        self.guard, self.seed, self.on, self.callback, self.mode, self.rng_seed = guard, seed, on, callback, mode, rng_seed
        _mode(mode)


class ChebyshevFilter(Solver):
    """``ChebyshevFilter(; tol=1e-8, degree=100, extra=40, maxiter=50, rescale_every=20,
    ortho=:svqb, bound=:gershgorin, seed=:product, on, callback)`` (src/tree.jl:293-321)."""

    def __init__(self, tol=1e-8, degree=100, extra=40, maxiter=50, rescale_every=20, ortho="svqb", bound="gershgorin",
                 seed="product", on=None, callback=None, mode="f64", rng_seed=0x4346):
        if not (tol > 0 and math.isfinite(tol)):
            raise L.ArgumentError(f"ChebyshevFilter: tol must be positive and finite, got {tol}")
        if not 1 <= degree <= 200:
            raise L.ArgumentError(f"ChebyshevFilter: degree must lie in 1:200, got {degree}")
        if extra < 2:
            raise L.ArgumentError("ChebyshevFilter: extra must be at least 2 so at least one column serves as a "
                                  f"convergence guard; the top column sets the filter cut and cannot be one, got {extra}")
        if maxiter < 1:
            raise L.ArgumentError(f"ChebyshevFilter: maxiter must be at least 1, got {maxiter}")
        if rescale_every < 1:
            raise L.ArgumentError(f"ChebyshevFilter: rescale_every must be at least 1, got {rescale_every}")
        if ortho not in ("svqb", "cholqr"):
            raise L.ArgumentError(f"ChebyshevFilter: ortho must be :svqb or :cholqr, got {ortho!r}")
        if not (bound == "gershgorin" or (isinstance(bound, (int, float)) and math.isfinite(bound))):
            raise L.ArgumentError(f"ChebyshevFilter: bound must be :gershgorin or a finite number, got {bound!r}")
        if seed not in ("product", "random"):
            raise L.ArgumentError(f"ChebyshevFilter: seed must be :product or :random, got {seed!r}")
        self.tol, self.degree, self.extra, self.maxiter = tol, degree, extra, maxiter
        self.rescale_every, self.ortho, self.bound, self.seed = rescale_every, ortho, bound, seed
        self.on, self.callback, self.mode, self.rng_seed = on, callback, mode, rng_seed
        _mode(mode)


class Auto(Solver):
    """``Auto(; dense_cutoff=1000, on)`` (src/tree.jl:343-351): Dense up to the
    cutoff (or when N < 3m+5), ChebyshevFilter at a device leaf, LOBPCG at a branch."""

    def __init__(self, dense_cutoff=1000, on=None, mode="f64"):
        if dense_cutoff < 0:
            raise L.ArgumentError(f"Auto: dense_cutoff must be nonnegative, got {dense_cutoff}")
        self.dense_cutoff, self.on, self.mode = dense_cutoff, on, mode


def resolve(s, info):
    """``resolve(::Auto, info)`` (src/materialize.jl:550-559); other solvers resolve to themselves."""
    if not isinstance(s, Auto):
        return s
    n = int(np.prod(info["dims"]))
    m = info["m"]
    if n <= s.dense_cutoff or (0 < m and n < 3 * m + 5):
        return Dense()
    if info["isleaf"]:
        return ChebyshevFilter(mode=s.mode)
    return LOBPCG(preconditioner="ritz", seed="product", mode=s.mode)


def _iter_info(ci):
    """The ``info`` a callback receives: LOBPCG's fields (n_iter, n_matvec, n_locked,
    λs of the active columns, their residual norms; src/lobpcg.jl:1-8) or
    ChebyshevFilter's (iteration, λ, residuals, n_locked, λ_cut, λ_max, degree;
    src/chebyshev.jl:362-363), copies the callback may keep."""
    import types
    w = ci.width
    lam = np.ctypeslib.as_array(ci.lambda_, shape=(w,)).copy() if w else np.zeros(0)
    res = np.ctypeslib.as_array(ci.residuals, shape=(w,)).copy() if w else np.zeros(0)
    prec = "f32" if ci.precision == L.KFC_PREC_F32 else "f64"
    if ci.solver == 0:
        a = ci.active0
        return types.SimpleNamespace(**{"n_iter": ci.iteration, "iteration": ci.iteration, "n_matvec": ci.n_matvec,
                                        "n_locked": ci.n_locked, "λs": lam[a:], "lambdas": lam, "residuals": res,
                                        "residual_norms": res[a:], "precision": prec})
    return types.SimpleNamespace(**{"iteration": ci.iteration, "λ": lam, "residuals": res, "n_locked": ci.n_locked,
                                    "λ_cut": ci.lambda_cut, "λ_max": ci.lambda_max, "degree": ci.degree,
                                    "n_matvec": ci.n_matvec, "precision": prec})


class _Callback:
    """The solver's ``callback`` installed on the operator's context for one solve
    (``kfc_ctx_set_callback``): called once per iteration with the info record, its
    return value ignored; an exception it raises stops the solve and propagates
    unchanged, with no retry (src/lobpcg.jl:247-253, 279)."""

    def __init__(self, H, user):
        self.H, self.user, self.exc, self.fn = H, user, None, None

    def __enter__(self):
        if self.user is not None:
            def trampoline(pinfo, _):
                try:
                    self.user(_iter_info(pinfo.contents))
                    return 0
                except BaseException as e:      # re-raised after the library returns
                    self.exc = e
                    return 1
            self.fn = L.CALLBACK(trampoline)
            L.check(L.lib().kfc_ctx_set_callback(self.H.ctx.handle, self.fn, None), self.H.ctx.handle)
        return self

    def __exit__(self, *a):
        if self.fn is not None:
            L.lib().kfc_ctx_set_callback(self.H.ctx.handle, L.CALLBACK(), None)
            self.fn = None
        return False

    def check(self, rc):
        if self.exc is not None:
            raise self.exc
        L.check(rc, self.H.ctx.handle)


def _out(H, m):
    E = np.empty(m)
    V = np.empty(H.N * m * (2 if H.dtype == np.complex128 else 1))
    return E, V


def solve_stats(s, H, m, X0=None):
    """``solve(s, H, m; X0)`` plus the solver's work counters."""
    lib = L.lib()
    st = L.Stats()
    if isinstance(s, Dense):
        k = m if 0 < m < H.N else H.N
        E, V = _out(H, k)
        L.check(lib.kfc_dense(H.handle, int(m), E.ctypes.data_as(L._D), V.ctypes.data_as(L._D), C.byref(st)),
                H.ctx.handle)
        return E, unpack_block(V, H.N, k, H.dtype), st.as_dict()
    E, V = _out(H, m)
    if isinstance(s, Lanczos):
        rc = lib.kfc_lanczos(H.handle, int(m), int(s.krylovdim), float(s.tol), int(s.maxiter), int(s.seed),
                             _mode(s.mode), int(s.verbosity), E.ctypes.data_as(L._D), V.ctypes.data_as(L._D),
                             C.byref(st))
    elif isinstance(s, LOBPCG):
        x0 = None
        seed_kind = L.KFC_SEED[s.seed]
        if X0 is not None:
            X0 = np.asarray(X0)
            w = m + guard_count(s, H.N, m)
            if X0.shape == (H.N, w) and w > m:
                seed_kind = L.KFC_SEED["warm_guards"]   # the warm start carries the guard columns
            elif X0.shape != (H.N, m):
                raise L.DimensionMismatch(f"LOBPCG: X0 has size {X0.shape}, expected ({H.N}, {m})")
            x0 = pack_block(X0, H.dtype, "LOBPCG: the warm-start block X0")
        with _Callback(H, s.callback) as cb:
            rc = lib.kfc_lobpcg(H.handle, int(m), int(s.guard), float(s.tol), int(s.maxiter),
                                L.KFC_PRECOND[s.preconditioner], float(s.eta), seed_kind,
                                x0.ctypes.data_as(L._D) if x0 is not None else None, int(s.rng_seed), _mode(s.mode),
                                E.ctypes.data_as(L._D), V.ctypes.data_as(L._D), C.byref(st))
        cb.check(rc)
    elif isinstance(s, ChebyshevFilter):
        with _Callback(H, s.callback) as cb:
            rc = lib.kfc_chebyshev(H.handle, int(m), int(s.degree), int(s.extra), int(s.maxiter),
                                   int(s.rescale_every), 0 if s.ortho == "svqb" else 1,
                                   math.nan if s.bound == "gershgorin" else float(s.bound), L.KFC_SEED[s.seed],
                                   float(s.tol), int(s.rng_seed), _mode(s.mode), E.ctypes.data_as(L._D),
                                   V.ctypes.data_as(L._D), C.byref(st))
        cb.check(rc)
    else:
        raise L.ArgumentError(f"no device method for solver {type(s).__name__}")
    L.check(rc, H.ctx.handle)
    return E, unpack_block(V, H.N, m, H.dtype), st.as_dict()


def guard_count(s, N, m):
    """Guard columns of a LOBPCG solve: min(guard, (N - 5) / 3 - m) (src/lobpcg.jl:243)."""
    return max(0, min(int(s.guard), (int(N) - 5) // 3 - int(m)))


def solve(s, H, m, X0=None):
    """``solve(s, H, m; X0) -> (E, V)`` (src/solvers.jl:3-31)."""
    E, V, _ = solve_stats(s, H, m, X0)
    return E, V


def residuals(H, E, V):
    """Fresh FP64 true residuals ||H v_j - E_j v_j|| (src/solvers.jl:65-69)."""
    m = len(E)
    E = np.ascontiguousarray(E, dtype=np.float64)
    Vb = pack_block(V, H.dtype, "residuals: V")
    out = np.empty(m)
    L.check(L.lib().kfc_residuals(H.handle, m, E.ctypes.data_as(L._D), Vb.ctypes.data_as(L._D),
                                  out.ctypes.data_as(L._D)), H.ctx.handle)
    return out


def solve_batch(s, Hs, m, X0s=None, lockstep=False):
    """Independent LOBPCG solves of a parameter sweep.  lockstep=False
    (``kfc_lobpcg_batch``): operators built on different contexts
    (``kfc._lib.Context``) run their solves concurrently on the device.
    lockstep=True (``kfc_lobpcg_lockstep``): real operators of one size on one
    context advance together, each LOBPCG step one launch for the whole sweep.
    Returns lists of (E, V) and the stats."""
    if not isinstance(s, LOBPCG):
        raise L.ArgumentError("solve_batch serves LOBPCG sweeps")
    if s.callback is not None:
        raise L.UnsupportedError("solve_batch runs many solves at once and calls no per-iteration callback; "
                                 "solve the points one at a time with solve() to use one")
    n = len(Hs)
    ops = (C.c_void_p * n)(*[H.handle.value for H in Hs])
    E = np.empty(n * m)
    Vb = [np.empty(H.N * m * (2 if H.dtype == np.complex128 else 1)) for H in Hs]
    Vp = (L._D * n)(*[v.ctypes.data_as(L._D) for v in Vb])
    x0p, keep = None, []
    seed_kind = L.KFC_SEED[s.seed]
    if X0s is not None:
        x0p = (L._D * n)()
        widths = set()
        for i, (H, X0) in enumerate(zip(Hs, X0s)):
            if X0 is None:
                x0p[i] = L._D()
            else:
                X0 = np.asarray(X0)
                w = m + guard_count(s, H.N, m)
                if X0.shape not in ((H.N, m), (H.N, w)):
                    raise L.DimensionMismatch(f"LOBPCG: X0 of instance {i} has size {X0.shape}, expected ({H.N}, {m})")
                widths.add(X0.shape[1] > m)
                b = pack_block(X0, H.dtype, f"LOBPCG: the warm-start block X0 of instance {i}")
                keep.append(b)
                x0p[i] = b.ctypes.data_as(L._D)
        if len(widths) > 1:
            raise L.ArgumentError("solve_batch: every X0 carries the guard columns or none does")
        if widths == {True}:
            if any(X0 is None for X0 in X0s):
                raise L.ArgumentError("solve_batch: warm starts with guard columns need an X0 for every instance")
            seed_kind = L.KFC_SEED["warm_guards"]
    st = (L.Stats * n)()
    fn = L.lib().kfc_lobpcg_lockstep if lockstep else L.lib().kfc_lobpcg_batch
    L.check(fn(n, ops, int(m), int(s.guard), float(s.tol), int(s.maxiter), L.KFC_PRECOND[s.preconditioner],
               float(s.eta), seed_kind, x0p, int(s.rng_seed), _mode(s.mode), E.ctypes.data_as(L._D), Vp, st),
            Hs[0].ctx.handle)
    out = [(E[i * m:(i + 1) * m].copy(), unpack_block(Vb[i], Hs[i].N, m, Hs[i].dtype)) for i in range(n)]
    return out, [x.as_dict() for x in st]


def solve_batch_subspace(s, Hs, m, B, rowmaps, coeffs=None):
    """A lockstep LOBPCG sweep (``kfc_lobpcg_lockstep_subspace``) whose points start
    from a learned subspace instead of the product seed: point i starts from the
    lowest m + g Ritz vectors of H_i on span(P_i B C_i).  ``B`` is a device float64
    tensor (nf, nb) whose columns span a parameter-independent frame of the parent
    basis; ``rowmaps[i]`` (nf int64) places frame row f at parent row rowmaps[i][f]
    (-1: not kept at that point); ``coeffs`` is None (all nb directions) or a list of
    (nb, ns) host arrays.  The start stands in for X0 (src/lobpcg.jl:217); the
    acceptance and outputs are those of ``solve_batch(..., lockstep=True)``."""
    import torch
    if not isinstance(s, LOBPCG):
        raise L.ArgumentError("solve_batch_subspace serves LOBPCG sweeps")
    if s.callback is not None:
        raise L.UnsupportedError("solve_batch_subspace runs many solves at once and calls no per-iteration callback")
    n = len(Hs)
    if len(rowmaps) != n or (coeffs is not None and len(coeffs) != n):
        raise L.DimensionMismatch(f"{n} operators, {len(rowmaps)} row maps"
                                  + ("" if coeffs is None else f" and {len(coeffs)} coefficient blocks"))
    if not (isinstance(B, torch.Tensor) and B.is_cuda and B.dtype == torch.float64 and B.dim() == 2):
        raise L.ArgumentError("solve_batch_subspace: B must be a 2-D float64 device tensor")
    nf, nb = B.shape
    Bc = B.t().contiguous()            # column-major (nf x nb): row j of the transpose is column j
    maps = [np.ascontiguousarray(np.asarray(r, dtype=np.int64).reshape(-1)) for r in rowmaps]
    for r in maps:
        if r.shape[0] != nf:
            raise L.DimensionMismatch(f"a row map has {r.shape[0]} entries for a frame of {nf} rows")
    mp = (C.POINTER(C.c_int64) * n)(*[r.ctypes.data_as(C.POINTER(C.c_int64)) for r in maps])
    ns, cp, keep = nb, None, []
    if coeffs is not None:
        blocks = [np.asfortranarray(np.asarray(c, dtype=np.float64)) for c in coeffs]
        ns = blocks[0].shape[1]
        for b in blocks:
            if b.shape != (nb, ns):
                raise L.DimensionMismatch(f"coefficient block of size {b.shape}, expected ({nb}, {ns})")
        keep = [b.reshape(-1, order="F") for b in blocks]
        cp = (L._D * n)(*[k.ctypes.data_as(L._D) for k in keep])
    ops = (C.c_void_p * n)(*[H.handle.value for H in Hs])
    E = np.empty(n * m)
    Vb = [np.empty(H.N * m) for H in Hs]
    Vp = (L._D * n)(*[v.ctypes.data_as(L._D) for v in Vb])
    st = (L.Stats * n)()
    torch.cuda.synchronize()            # B is read on the library's stream
    L.check(L.lib().kfc_lobpcg_lockstep_subspace(
        n, ops, int(m), int(s.guard), float(s.tol), int(s.maxiter), L.KFC_PRECOND[s.preconditioner], float(s.eta),
        C.c_void_p(Bc.data_ptr()), int(nf), int(nb), mp, cp, int(ns), int(s.rng_seed), _mode(s.mode),
        E.ctypes.data_as(L._D), Vp, st), Hs[0].ctx.handle)
    del Bc
    out = [(E[i * m:(i + 1) * m].copy(), unpack_block(Vb[i], Hs[i].N, m, Hs[i].dtype)) for i in range(n)]
    return out, [x.as_dict() for x in st]


def solve_dense_batch(Hs, m):
    """``solve(Dense(), H, m)`` for every operator of a sweep at once
    (``kfc_dense_batch``): small operators' matrices formed and diagonalized on host
    threads.  Returns lists of (E, V) and the stats."""
    n = len(Hs)
    if n == 0:
        return [], []
    ops = (C.c_void_p * n)(*[H.handle.value for H in Hs])
    E = np.empty(n * m)
    Vb = [np.empty(H.N * m * (2 if H.dtype == np.complex128 else 1)) for H in Hs]
    Vp = (L._D * n)(*[v.ctypes.data_as(L._D) for v in Vb])
    st = (L.Stats * n)()
    L.check(L.lib().kfc_dense_batch(n, ops, int(m), E.ctypes.data_as(L._D), Vp, st), Hs[0].ctx.handle)
    out = [(E[i * m:(i + 1) * m].copy(), unpack_block(Vb[i], Hs[i].N, m, Hs[i].dtype)) for i in range(n)]
    return out, [x.as_dict() for x in st]


def solve_chebyshev_batch(Hs, m, extra=8, degree=60, tol=1e-9, maxiter=200, mode="mixed", seed=0x4346):
    """Batched ChebyshevFilter solves of one-axis operators on one context
    (``kfc_chebyshev_batch``): every filter step of the whole batch is one launch.
    Returns lists of (E, V) and the stats."""
    n = len(Hs)
    if n == 0:
        return [], []
    ops = (C.c_void_p * n)(*[H.handle.value for H in Hs])
    E = np.empty(n * m)
    Vb = [np.empty(H.N * m * (2 if H.dtype == np.complex128 else 1)) for H in Hs]
    Vp = (L._D * n)(*[v.ctypes.data_as(L._D) for v in Vb])
    st = (L.Stats * n)()
    L.check(L.lib().kfc_chebyshev_batch(n, ops, int(m), int(extra), int(degree), float(tol), int(maxiter), int(seed),
                                        _mode(mode), E.ctypes.data_as(L._D), Vp, st), Hs[0].ctx.handle)
    out = [(E[i * m:(i + 1) * m].copy(), unpack_block(Vb[i], Hs[i].N, m, Hs[i].dtype)) for i in range(n)]
    return out, [x.as_dict() for x in st]
