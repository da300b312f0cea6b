"""The resolvent (s - H)^-1 of a node applied away from chosen kept states, with the
linear solve on the orthogonal complement of the known states on the device: the
Sternheimer step of the derivatives (src/resolvent.jl).

``complement_solve`` is ``_complement_solve`` (src/resolvent.jl:146-256) over the C ABI
(``kfc_complement_solve``); ``Resolvent`` / ``resolvent`` / ``apply_resolvent`` mirror
``Resolvent``, ``_resolvent`` and ``_apply_resolvent`` (:13-118, 270-331) for the nodes
this package materializes (kept states only: a node keeps its lowest states, so the
complement of the kept ones is an invariant subspace above them).

Blocks are device tensors in the layout of ``KronFreeOperator.mul``: a (n, N) tensor
whose rows are the columns of the N x n block."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L

MAX_COLUMNS = 64   # columns per lockstep solve (the C entry point's limit)


def _dev(H, A, name):
    import torch
    dt = torch.complex128 if H.dtype == np.complex128 else torch.float64
    t = A if isinstance(A, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(A))
    t = t.to(device=f"cuda:{H.ctx.device}", dtype=dt).contiguous()
    if t.shape[-1] != H.N:
        raise L.DimensionMismatch(f"{name}: rows of {t.shape[-1]} entries for an operator of size {H.N}")
    return t


def complement_solve(H, D, basis, B, shifts, tol, maxiter, eta=0.01, scale=None):
    """``_complement_solve(H, D, basis, B, shifts; tol, maxiter, eta, scale)``
    (src/resolvent.jl:146-256): row j of X solves Q (H - shifts[j]) Q x = Q B[j] with x
    orthogonal to the rows of ``basis`` (orthonormal).  ``D`` is H's real diagonal
    (device float64), ``basis`` (k, N), ``B`` (n, N).  Returns (X, iterations, residuals,
    from_start, at_floor, floor_residuals) as the reference does, indices 0-based; X is
    a device tensor.  More than 64 columns are solved 64 at a time."""
    import torch
    B = _dev(H, B, "complement_solve")
    if B.dim() == 1:
        B = B[None, :]
    n = B.shape[0]
    basis = _dev(H, basis, "complement_solve") if basis is not None and len(basis) else None
    k = 0 if basis is None else basis.shape[0]
    D = torch.as_tensor(D).to(device=B.device, dtype=torch.float64).contiguous().reshape(-1)
    if D.numel() != H.N:
        raise L.DimensionMismatch(f"complement_solve: the diagonal has {D.numel()} entries for an operator of "
                                  f"size {H.N}")
    shifts = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(-1))
    if len(shifts) != n:
        raise L.DimensionMismatch(f"{n} columns and {len(shifts)} shifts")
    if scale is not None:
        scale = np.ascontiguousarray(np.asarray(scale, dtype=np.float64).reshape(-1))
        if len(scale) != n:
            raise L.DimensionMismatch(f"{n} columns and {len(scale)} scales")
    X = torch.empty_like(B)
    its = 0
    res, flags, flr = np.zeros(n), np.zeros(n, dtype=np.int32), np.zeros(n)
    torch.cuda.synchronize()
    for c0 in range(0, n, MAX_COLUMNS):
        c1 = min(n, c0 + MAX_COLUMNS)
        it = C.c_int64()
        sc = None if scale is None else scale[c0:c1].ctypes.data_as(L._D)
        L.check(L.lib().kfc_complement_solve(
            H.handle, C.c_void_p(D.data_ptr()), C.c_void_p(basis.data_ptr() if k else 0), k,
            C.c_void_p(B[c0].data_ptr()), c1 - c0, shifts[c0:c1].ctypes.data_as(L._D), float(tol), int(maxiter),
            float(eta), sc, C.c_void_p(X[c0].data_ptr()), C.byref(it), res[c0:c1].ctypes.data_as(L._D),
            flags[c0:c1].ctypes.data_as(C.POINTER(C.c_int32)), flr[c0:c1].ctypes.data_as(L._D)), H.ctx.handle)
        its = max(its, it.value)
    H.ctx.sync()
    from_start = [j for j in range(n) if flags[j] & 2]
    at_floor = [j for j in range(n) if flags[j] & 4]
    return X, its, res, from_start, at_floor, flr


@dataclass
class Complement:
    """``_Complement`` (src/resolvent.jl:33-41): the node's operator, its real diagonal
    (device) and the known states (device rows), with the solver limits."""
    H: object
    D: object
    basis: object
    tol: float
    maxiter: int
    log: list = None


@dataclass
class Resolvent:
    """``Resolvent`` (src/resolvent.jl:13-21): kept energies and vectors (host, N x k),
    the other known states (none here), and the complement solve (or None when the
    kept states span the space)."""
    energies: np.ndarray
    vectors: np.ndarray
    complement: Complement
    degenerate_tol: float


def resolvent(mn, degenerate_tol=1e-9, lintol=1e-10, linmaxiter=5000, log=None):
    """``_resolvent(mn, H; degenerate_tol, lintol, linmaxiter, log)`` (src/resolvent.jl:51-91)
    for a materialized node that keeps its lowest states: the complement solve runs on
    the node's operator when the kept states do not span its space."""
    V = np.asarray(mn.vectors)
    N, k = V.shape
    if k == N:
        return Resolvent(np.array(mn.energies), V, None, float(degenerate_tol))
    H = mn.operator
    if H is None:
        raise L.ArgumentError("the node does not hold its complete spectrum; its forward operator is needed for the "
                              "linear solve on the complement of its computed states")
    c = Complement(H, H.diagonal_device(), _dev(H, V.T, "resolvent"), float(lintol), int(linmaxiter), log)
    return Resolvent(np.array(mn.energies), V, c, float(degenerate_tol))


def apply_resolvent(res, R, shifts, exclude):
    """``_apply_resolvent(res, R, shifts, exclude)`` (src/resolvent.jl:270-331):
    W[:, j] = sum_l psi_l <psi_l|R[:, j]> / (shifts[j] - E_l) over every state of the
    node's space except the kept states listed in exclude[j] (0-based), the part on the
    complement of the kept states from the complement solve of -Q R[:, j].  R is a host
    N x n array.  Returns (W, clashes): a clashing column (a kept state within
    degenerate_tol of the shift with a nonzero overlap; 0 marks none here) is NaN, as
    is a column of R that is not finite."""
    R = np.array(R, dtype=np.result_type(res.vectors.dtype, np.asarray(R).dtype), copy=True)
    if R.ndim == 1:
        R = R[:, None]
    n = R.shape[1]
    shifts = np.asarray(shifts, dtype=float)
    if not (len(shifts) == len(exclude) == n):
        raise L.DimensionMismatch(f"{n} columns, {len(shifts)} shifts and {len(exclude)} exclusion lists")
    nonfinite = [j for j in range(n) if not np.all(np.isfinite(R[:, j]))]
    R[:, nonfinite] = 0
    Ck = res.vectors.conj().T @ R
    Rc = R - res.vectors @ Ck if res.complement is not None else None
    clashes = [[] for _ in range(n)]
    for j in range(n):
        zero_overlap = 1e-10 * np.linalg.norm(R[:, j])
        for l in range(Ck.shape[0]):
            if l in exclude[j]:
                Ck[l, j] = 0
                continue
            d = shifts[j] - res.energies[l]
            if abs(d) <= res.degenerate_tol:
                if abs(Ck[l, j]) > zero_overlap:
                    clashes[j].append(l)
                Ck[l, j] = 0
            else:
                Ck[l, j] /= d
    W = res.vectors @ Ck
    c = res.complement
    if c is not None:
        scale = np.linalg.norm(R, axis=0)
        X, iterations, residuals, from_start, at_floor, floor_residuals = complement_solve(
            c.H, c.D, c.basis, (-Rc).T, shifts, c.tol, c.maxiter, scale=scale)
        W = W + X.cpu().numpy().T
        if c.log is not None:
            c.log.append(dict(columns=n, iterations=iterations, residual=float(np.max(residuals, initial=0.0)),
                              from_start=from_start, at_floor=at_floor,
                              floor_residual=float(np.max(floor_residuals, initial=0.0))))
    for j in range(n):
        if clashes[j] or j in nonfinite:
            W[:, j] = np.nan
    return W, clashes
