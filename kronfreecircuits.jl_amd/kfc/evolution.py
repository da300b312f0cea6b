"""Matrix-free time evolution on a node's product basis: the consumers of
``port_operator`` (src/portoperator.jl:22-39) in the QuantumToolbox extension
(ext/KronFreeCircuitsQuantumToolboxExt.jl), on the device.

``KronFreeApply`` and ``InteractionApply`` are the two applies the extension hands
to QuantumToolbox's ODE solvers (``kfc_apply_affine``, ``kfc_apply_interaction``);
``DriveHamiltonian`` assembles them as ``drive_hamiltonian(sys, drives, node;
basis=:product, frame)`` does; ``evolve_rk4`` is a fixed-step integrator of
``dψ/dt = -i H(t) ψ`` over those applies (QuantumToolbox itself is Julia and out of
reach here), used to check the frames against each other and against exp(-iHt).

Kets are complex128 device tensors: an N-vector or a (b, N) block whose rows are the
columns of the N x b block (the layout of ``KronFreeOperator.mul``)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .tree import port_operator


def _block(H, V, name):
    import torch
    if V.dtype != torch.complex128 or not V.is_cuda:
        raise L.ArgumentError(f"{name}: kets are complex128 device tensors")
    if V.shape[-1] != H.N:
        raise L.DimensionMismatch(f"{name}: a ket of {V.shape[-1]} entries for an operator of size {H.N}")
    if V.stride(-1) != 1:
        raise L.ArgumentError(f"{name}: blocks must be row-contiguous")
    if V.dim() == 1:
        return 1, H.N
    if V.dim() != 2:
        raise L.DimensionMismatch(f"{name}: a block is a (b, N) tensor or an N-vector")
    return V.shape[0], V.stride(0)


def _check_pair(H, w, v, name):
    if w.shape != v.shape:
        raise L.DimensionMismatch(f"{name}: size(w) = {tuple(w.shape)} differs from size(v) = {tuple(v.shape)}")
    b, ld = _block(H, v, name)
    bw, ldw = _block(H, w, name)
    if ld != ldw:
        raise L.ArgumentError(f"{name}: w and v must have equal row strides")
    return b, ld


class KronFreeApply:
    """``KronFreeApply(H, scale, shift)`` (ext/KronFreeCircuitsQuantumToolboxExt.jl:44-87):
    ``A(w, v, alpha=1, beta=0)`` sets ``w = alpha scale (H - shift) v + beta w``
    (``beta == 0`` leaves w unread) and returns w."""

    def __init__(self, H, scale=1.0, shift=0.0):
        self.H, self.scale, self.shift = H, float(scale), float(shift)

    def __call__(self, w, v, alpha=1.0, beta=0.0):
        import torch
        b, ld = _check_pair(self.H, w, v, "KronFreeApply")
        a, bb = complex(alpha), complex(beta)
        torch.cuda.synchronize()
        L.check(L.lib().kfc_apply_affine(self.H.handle, C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), b, ld,
                                         a.real, a.imag, self.scale, self.shift, bb.real, bb.imag), self.H.ctx.handle)
        self.H.ctx.sync()
        return w


class InteractionApply:
    """``InteractionApply`` (ext/KronFreeCircuitsQuantumToolboxExt.jl:211-251): in the
    frame of the diagonal Omega = 2 pi (real(diag H0) - E0),
    ``A(w, v, t, p=None, alpha=1, beta=0)`` sets ``w = alpha e^{i Omega t} [2 pi (H0 -
    E0) - diag Omega + sum_k f_k(p, t) 2 pi O_k] e^{-i Omega t} v + beta w``."""

    def __init__(self, H0, E0, drives):
        self.H0, self.E0 = H0, float(E0)
        self.drives = list(drives)   # (operator, f(p, t))
        for O, _ in self.drives:
            if O.N != H0.N:
                raise L.DimensionMismatch("InteractionApply: a drive on another product basis")
        self.omega = (H0.diagonal_device() - self.E0) * (2 * math.pi)
        self._ops = (C.c_void_p * max(1, len(self.drives)))(*[O.handle.value for O, _ in self.drives])

    def __call__(self, w, v, t, p=None, alpha=1.0, beta=0.0):
        import torch
        b, ld = _check_pair(self.H0, w, v, "InteractionApply")
        f = np.zeros(2 * max(1, len(self.drives)))
        for k, (_, fk) in enumerate(self.drives):
            z = complex(fk(p, t))
            f[2 * k], f[2 * k + 1] = z.real, z.imag
        a, bb = complex(alpha), complex(beta)
        torch.cuda.synchronize()
        L.check(L.lib().kfc_apply_interaction(self.H0.handle, self.E0, C.c_void_p(self.omega.data_ptr()), float(t),
                                              self._ops, f.ctypes.data_as(L._D), len(self.drives),
                                              C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), b, ld,
                                              a.real, a.imag, bb.real, bb.imag), self.H0.ctx.handle)
        self.H0.ctx.sync()
        return w


class DriveHamiltonian:
    """``drive_hamiltonian(sys, drives, node; basis=:product, frame)``
    (ext/KronFreeCircuitsQuantumToolboxExt.jl:185-212, 253-280) on the product basis of
    the materialized node ``mn``: H(t) = 2 pi (H0 - E0) + sum_k f_k(p, t) 2 pi O_k with
    H0 the node's operator, E0 its lowest energy and O_k = port_operator(mn, key_k).
    ``drives`` is a list of ``(key, f)``.  ``frame="interaction"`` rotates by the
    diagonal Omega = 2 pi (real(diag H0) - E0); ``to_lab`` maps a state back.
    ``apply(w, v, t, p=None, alpha=1, beta=0)`` sets ``w = alpha H(t) v + beta w``
    (the interaction-frame H_I(t) with that frame)."""

    def __init__(self, mn, drives, frame="lab"):
        if frame not in ("lab", "interaction"):
            raise L.ArgumentError(f"frame must be lab or interaction, got {frame}")
        for d in drives:
            if not (isinstance(d, tuple) and len(d) == 2 and callable(d[1])):
                raise L.ArgumentError(f"each drive must be (key, f) with f(p, t) a coefficient, got {d}")
        if mn.operator is None:
            raise L.ArgumentError("DriveHamiltonian: the node keeps no operator")
        self.mn, self.frame = mn, frame
        self.H0 = mn.operator
        self.E0 = float(mn.energies[0])
        self.ops = [(port_operator(mn, key, ctx=self.H0.ctx), f) for key, f in drives]
        if frame == "lab":
            self._h0 = KronFreeApply(self.H0, 2 * math.pi, self.E0)
            self._dr = [(KronFreeApply(O, 2 * math.pi, 0.0), f) for O, f in self.ops]
        else:
            self._ia = InteractionApply(self.H0, self.E0, self.ops)

    @property
    def omega(self):
        return self._ia.omega if self.frame == "interaction" else None

    def apply(self, w, v, t, p=None, alpha=1.0, beta=0.0):
        if self.frame == "interaction":
            return self._ia(w, v, t, p, alpha, beta)
        self._h0(w, v, alpha, beta)
        for A, f in self._dr:
            A(w, v, complex(alpha) * complex(f(p, t)), 1.0)
        return w

    def to_lab(self, psi, t):
        """``to_lab(frame, ψ, t)``: e^{-i Omega t} ψ in the interaction frame, ψ in the lab frame."""
        import torch
        if self.frame == "lab":
            return psi
        return torch.exp(-1j * self._ia.omega * t).to(psi.dtype) * psi


def evolve_rk4(Hd, psi0, t0, t1, steps, p=None):
    """ψ(t1) from ψ(t0) = psi0 under dψ/dt = -i H(t) ψ, classical fourth-order
    Runge-Kutta with `steps` equal steps; H(t) applied by ``Hd.apply``."""
    import torch
    psi = psi0.clone()
    h = (t1 - t0) / steps
    k = [torch.empty_like(psi) for _ in range(4)]
    tmp = torch.empty_like(psi)
    for n in range(steps):
        t = t0 + n * h
        Hd.apply(k[0], psi, t, p, alpha=-1j)
        torch.add(psi, k[0], alpha=h / 2, out=tmp)
        Hd.apply(k[1], tmp, t + h / 2, p, alpha=-1j)
        torch.add(psi, k[1], alpha=h / 2, out=tmp)
        Hd.apply(k[2], tmp, t + h / 2, p, alpha=-1j)
        torch.add(psi, k[2], alpha=h, out=tmp)
        Hd.apply(k[3], tmp, t + h, p, alpha=-1j)
        psi += (h / 6) * (k[0] + 2 * k[1] + 2 * k[2] + k[3])
    return psi
