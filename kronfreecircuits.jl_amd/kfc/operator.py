"""``AxisTerm`` and ``KronFreeOperator`` on the MI355X (src/apply.jl:157-253, 416-435).

The operator's data live on the device from construction; ``mul`` applies it
to device blocks (torch tensors) through ``kfc_apply``.  A block of ``b``
columns is a torch tensor of shape ``(b, N)`` (each row one column of the
reference's column-major ``N x b`` block) or a vector of length ``N``."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .localops import BandedOp, DenseOp, DiagonalOp, LocalOp, ShiftOp, is_complex


class AxisTerm:
    """``AxisTerm(coeff, [(axis, op), ...])``, axes 0-based and strictly
    increasing (src/apply.jl:157-166)."""

    def __init__(self, coeff, factors):
        axes = [a for a, _ in factors]
        if any(b <= a for a, b in zip(axes, axes[1:])):
            raise L.ArgumentError(f"AxisTerm: factor axes must be strictly increasing, got {axes}")
        for _, op in factors:
            if not isinstance(op, LocalOp):
                raise L.ArgumentError("AxisTerm: factors must be (axis, LocalOp) pairs")
        self.coeff = coeff
        self.factors = list(factors)


def _cplx_value(x):
    return isinstance(x, complex) or np.iscomplexobj(x)


def _f64(a):
    a = np.ascontiguousarray(a)
    if np.iscomplexobj(a):
        return np.ascontiguousarray(np.stack([a.real, a.imag], axis=-1).reshape(-1)).astype(np.float64), 1
    return a.astype(np.float64).reshape(-1), 0


class KronFreeOperator:
    """``KronFreeOperator(dims, terms; constant=0, block_capacity=1, hermitian=true)``
    (src/apply.jl:222-253): H = sum_t c_t (x)_k A_{t,k} + constant I, axis 0
    fastest.  All-diagonal terms and the constant are fused into the diagonal
    by the library; the element type is complex when any coefficient, the
    constant or any factor's data is complex (src/apply.jl:232-233)."""

    def __init__(self, dims, terms, constant=0.0, block_capacity=1, hermitian=True, device=0, ctx=None):
        dims = [int(d) for d in dims]
        if any(d < 1 for d in dims):
            raise L.ArgumentError(f"KronFreeOperator: dims must be positive, got {dims}")
        for t in terms:
            for a, op in t.factors:
                if not 0 <= a < len(dims):
                    raise L.ArgumentError(f"KronFreeOperator: factor axis {a} outside 0:{len(dims) - 1}")
                if op.n != dims[a]:
                    raise L.DimensionMismatch(f"KronFreeOperator: factor on axis {a} has dimension {op.n}, "
                                              f"expected {dims[a]}")
        if block_capacity < 1:
            raise L.ArgumentError("KronFreeOperator: block_capacity must be positive")
        cplx = _cplx_value(constant) or any(_cplx_value(t.coeff) for t in terms) or \
            any(is_complex(op) for t in terms for _, op in t.factors)
        self.dims = dims
        self.N = int(np.prod(dims))
        self.dtype = np.complex128 if cplx else np.float64
        self.terms = terms
        self.constant = constant
        self.hermitian = hermitian
        self.ctx = ctx if ctx is not None else L.context(device)
        keep = []
        tdescs = (L.TermDesc * max(1, len(terms)))()
        for ti, t in enumerate(terms):
            fds = (L.FactorDesc * max(1, len(t.factors)))()
            for k, (a, op) in enumerate(t.factors):
                f = fds[k]
                f.axis, f.d = a, op.n
                if isinstance(op, DiagonalOp):
                    arr, c = _f64(op.d)
                    f.kind, f.is_complex, f.diag = L.KFC_DIAG, c, arr.ctypes.data_as(L._D)
                    keep.append(arr)
                elif isinstance(op, DenseOp):
                    arr, c = _f64(np.asfortranarray(op.M).reshape(-1, order="F"))
                    f.kind, f.is_complex, f.matrix = L.KFC_DENSE, c, arr.ctypes.data_as(L._D)
                    keep.append(arr)
                elif isinstance(op, BandedOp):
                    W = np.stack(op.weights, axis=1) if op.weights else np.zeros((op.n, 0))
                    arr, c = _f64(W.reshape(-1, order="F"))
                    offs = np.asarray(op.offsets, dtype=np.int64)
                    f.kind, f.is_complex, f.nbands = L.KFC_BANDED, c, len(op.offsets)
                    f.weights = arr.ctypes.data_as(L._D)
                    f.offsets = offs.ctypes.data_as(C.POINTER(C.c_int64))
                    keep += [arr, offs]
                elif isinstance(op, ShiftOp):
                    fc = complex(op.factor)
                    f.kind, f.shift, f.is_complex = L.KFC_SHIFT, op.s, 1
                    f.factor_re, f.factor_im = fc.real, fc.imag
                else:
                    raise L.ArgumentError(f"unknown local operator {type(op).__name__}")
            keep.append(fds)
            c = complex(t.coeff)
            tdescs[ti].coeff_re, tdescs[ti].coeff_im = c.real, c.imag
            tdescs[ti].nfactors = len(t.factors)
            tdescs[ti].factors = fds
        dims_arr = np.asarray(dims, dtype=np.int64)
        desc = L.OpDesc()
        desc.ndims = len(dims)
        desc.dims = dims_arr.ctypes.data_as(C.POINTER(C.c_int64))
        desc.dtype = L.KFC_C128 if cplx else L.KFC_F64
        cc = complex(constant)
        desc.constant_re, desc.constant_im = cc.real, cc.imag
        desc.nterms = len(terms)
        desc.terms = tdescs
        desc.block_capacity = int(block_capacity)
        desc.hermitian = 1 if hermitian else 0
        h = C.c_void_p()
        L.check(L.lib().kfc_op_create(self.ctx.handle, C.byref(desc), C.byref(h)), self.ctx.handle)
        self.handle = h

    def __del__(self):
        try:
            if getattr(self, "handle", None) and L._lib is not None:
                L._lib.kfc_op_destroy(self.handle)
        except Exception:
            pass

    @property
    def fused(self):
        """True when every application runs as one fused stencil pass."""
        return bool(L.lib().kfc_op_fused(self.handle))

    @property
    def stencil(self):
        """The fused stencil's launch shape ``(variant, march)`` (the autotuner's
        pick after the first application)."""
        v, m = C.c_int32(), C.c_int32()
        L.check(L.lib().kfc_op_stencil_get(self.handle, C.byref(v), C.byref(m)), self.ctx.handle)
        return v.value, m.value

    def work(self):
        """``(bytes, flops)`` of one column of ``mul``: the roofline numerators."""
        b, f = C.c_double(), C.c_double()
        L.check(L.lib().kfc_op_work(self.handle, C.byref(b), C.byref(f)), self.ctx.handle)
        return b.value, f.value

    def fill_scratch(self, value):
        """Fill the library's internal scratch of this operator with ``value`` (the
        reference's device tests fill ``H.buf1``/``H.buf2`` with NaN or 1e300 to show
        that no result reads stale scratch, test/cuda/test_device.jl:163-207)."""
        L.check(L.lib().kfc_op_fill_scratch(self.handle, float(value)), self.ctx.handle)

    def set_stencil(self, variant, march=0):
        """Pin the fused stencil's launch shape; ``variant=-1`` hands it back to the autotuner."""
        L.check(L.lib().kfc_op_stencil_set(self.handle, int(variant), int(march)), self.ctx.handle)

    @property
    def shape(self):
        return (self.N, self.N)

    def _blk(self, X):
        if X.dim() == 1:
            return 1, self.N
        if X.dim() != 2:
            raise L.DimensionMismatch("a block is a (b, N) tensor or an N-vector")
        return X.shape[0], X.stride(0)

    def _check_tensors(self, name, prec, *tensors):
        """Device tensors of this operator's device whose element type matches the
        operator's type at precision ``prec`` (the reference dispatches on the element
        type and refuses what it cannot apply)."""
        import torch
        if prec not in ("f64", "f32"):
            raise L.ArgumentError(f"{name}: precision must be 'f64' or 'f32', got {prec!r}")
        cplx = self.dtype == np.complex128
        want = {("f64", False): torch.float64, ("f32", False): torch.float32,
                ("f64", True): torch.complex128, ("f32", True): torch.complex64}[(prec, cplx)]
        for T in tensors:
            if not T.is_cuda or (T.device.index or 0) != self.ctx.device:
                raise L.ArgumentError(f"{name}: blocks must live on cuda:{self.ctx.device}, got {T.device}")
            if T.dtype != want:
                raise L.ArgumentError(f"{name}: a {'complex' if cplx else 'real'} operator at precision {prec!r} "
                                      f"takes {want} blocks, got {T.dtype}")

    def mul(self, Y, X, precision="f64"):
        """``mul!(Y, H, X)`` (src/apply.jl:416-432) on device tensors.  Complex128
        blocks on a real operator take ``_mul_real_operator!`` (src/apply.jl:397-407):
        H applied to the real and imaginary parts.  The library's stream is ordered
        after torch's current stream before the launch, and torch's current stream
        after the launch, so torch tensors go in and out with no explicit sync."""
        import torch
        if X.shape != Y.shape:
            raise L.DimensionMismatch(f"mul!: size(y) = {tuple(Y.shape)} differs from size(x) = {tuple(X.shape)}")
        if X.shape[-1] != self.N:
            raise L.DimensionMismatch(f"KronFreeOperator of size {self.shape} applied to x with {X.shape[-1]} rows")
        if X.stride(-1) != 1 or Y.stride(-1) != 1 or (X.dim() == 2 and X.stride(0) != Y.stride(0)):
            raise L.ArgumentError("mul!: blocks must be row-contiguous with equal row strides")
        if self.dtype == np.float64 and X.dtype == torch.complex128 and precision == "f64":
            self._check_cplx_on_real(X, Y)
            prec = L.KFC_PREC_Z64
        else:
            self._check_tensors("mul!", precision, X, Y)
            prec = L.KFC_PREC_F64 if precision == "f64" else L.KFC_PREC_F32
        b, ld = self._blk(X)
        self.ctx.wait_torch()
        L.check(L.lib().kfc_apply(self.handle, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), b, ld, prec),
                self.ctx.handle)
        self.ctx.torch_wait()
        return Y

    def _check_cplx_on_real(self, X, Y):
        import torch
        for T in (X, Y):
            if not T.is_cuda or (T.device.index or 0) != self.ctx.device:
                raise L.ArgumentError(f"mul!: blocks must live on cuda:{self.ctx.device}, got {T.device}")
            if T.dtype != torch.complex128:
                raise L.ArgumentError(f"mul!: a real operator takes complex128 blocks only as complex128 in and out, "
                                      f"got {T.dtype}")

    def cheb_step_dots(self, Y, B, A=None, a=1.0, b=0.0, g=0.0, precision="f64"):
        """``Y = a H B + b B + g A`` on one device vector (the step of
        ``_chebyshev_step!``, src/chebyshev.jl:25-30) with ``(B.B, B.Y)`` as FP64
        sums.  Returns (dots, fused): fused is True when the fused plane sweep
        produced the sums in the apply's own pass."""
        for T in (B, Y) + ((A,) if A is not None else ()):
            if T.dim() != 1 or T.shape[0] != self.N or T.stride(0) != 1:
                raise L.DimensionMismatch(f"cheb_step_dots: contiguous vectors of {self.N} entries")
        self._check_tensors("cheb_step_dots", precision, B, Y, *((A,) if A is not None else ()))
        prec = L.KFC_PREC_F64 if precision == "f64" else L.KFC_PREC_F32
        d = np.zeros(2)
        fz = C.c_int32(0)
        self.ctx.wait_torch()
        L.check(L.lib().kfc_apply_cheb_dots(self.handle, C.c_void_p(B.data_ptr()),
                                            C.c_void_p(A.data_ptr() if A is not None else 0),
                                            C.c_void_p(Y.data_ptr()), float(a), float(b), float(g), prec,
                                            d.ctypes.data_as(L._D), C.byref(fz)), self.ctx.handle)
        self.ctx.torch_wait()
        return d, bool(fz.value)

    def __matmul__(self, X):
        """``H * x`` (src/apply.jl:434-435): a new block of the promoted element type."""
        import torch
        Y = torch.empty_like(X)
        self.mul(Y, X, "f64" if X.dtype in (torch.float64, torch.complex128) else "f32")
        return Y

    def sync(self):
        self.ctx.sync()

    # -- derived quantities (src/apply.jl:455-685), host data ---------------------------
    def diagonal(self):
        """``_full_diagonal(H)``: the real diagonal of H."""
        out = np.empty(self.N)
        L.check(L.lib().kfc_op_diagonal(self.handle, out.ctypes.data_as(L._D)), self.ctx.handle)
        return out

    def gershgorin_bound(self):
        out = C.c_double()
        L.check(L.lib().kfc_op_gershgorin(self.handle, C.byref(out)), self.ctx.handle)
        return out.value

    def product_seed(self, k):
        buf = np.empty(self.N * k * (2 if self.dtype == np.complex128 else 1))
        L.check(L.lib().kfc_op_product_seed(self.handle, int(k), buf.ctypes.data_as(L._D)), self.ctx.handle)
        return unpack_block(buf, self.N, k, self.dtype)

    def diagonal_device(self):
        """``_full_diagonal(H)`` formed on the device (kfc_op_diagonal_device): a torch
        float64 tensor of length N on this operator's device."""
        import torch
        D = torch.empty(self.N, dtype=torch.float64, device=f"cuda:{self.ctx.device}")
        self.ctx.wait_torch()
        L.check(L.lib().kfc_op_diagonal_device(self.handle, C.c_void_p(D.data_ptr())), self.ctx.handle)
        return D

    def product_seed_device(self, k):
        """``product_seed(H, k)`` formed on the device (kfc_op_product_seed_device): an
        N x k torch tensor (column-major storage, float64 or complex128) on this
        operator's device."""
        import torch
        dt = torch.complex128 if self.dtype == np.complex128 else torch.float64
        X = torch.empty(int(k), self.N, dtype=dt, device=f"cuda:{self.ctx.device}")
        self.ctx.wait_torch()
        L.check(L.lib().kfc_op_product_seed_device(self.handle, int(k), C.c_void_p(X.data_ptr()), self.N),
                self.ctx.handle)
        return X.t()


def unpack_block(buf, N, k, dtype):
    if dtype == np.complex128:
        z = buf.reshape(-1, 2)
        return (z[:, 0] + 1j * z[:, 1]).reshape(N, k, order="F")
    return buf.reshape(N, k, order="F")


def pack_block(X, dtype, name="block"):
    """A host block in the C ABI's layout (column-major, complex interleaved).  A
    complex block for a real operator is accepted only when it is exactly real
    (src/lobpcg.jl:226-230); anything else raises instead of losing its imaginary part."""
    X = np.asarray(X)
    if dtype == np.complex128:
        Xf = np.asfortranarray(X.astype(np.complex128)).reshape(-1, order="F")
        return np.ascontiguousarray(np.stack([Xf.real, Xf.imag], axis=-1).reshape(-1))
    if np.iscomplexobj(X):
        if np.any(np.imag(X) != 0):
            raise L.ArgumentError(f"{name} is complex but the operator is real; pass a real block or complexify "
                                  "the operator")
        X = np.real(X)
    return np.ascontiguousarray(np.asfortranarray(X.astype(np.float64)).reshape(-1, order="F"))
