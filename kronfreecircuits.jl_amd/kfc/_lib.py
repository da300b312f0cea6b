"""ctypes binding of libkfc.so (include/kfc.h).

The library is loaded from this directory; a missing library raises at import
of any entry point that needs it (there is no CPU fallback)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkfc.so")

KFC_OK, KFC_EARG, KFC_EDIM, KFC_ENOCONV, KFC_EBREAKDOWN, KFC_EDEVICE, KFC_ENOMEM, KFC_EUNSUPPORTED, KFC_EABORT = range(9)
KFC_F64, KFC_C128 = 0, 1
KFC_PREC_F64, KFC_PREC_F32, KFC_PREC_Z64 = 0, 1, 2
KFC_MODE_F64, KFC_MODE_MIXED = 0, 1
KFC_DIAG, KFC_BANDED, KFC_SHIFT, KFC_DENSE = 0, 1, 2, 3
KFC_PRECOND = {"ritz": 0, "jacobi": 1, "none": 2}
KFC_SEED = {"product": 0, "random": 1, "warm_guards": 2}


class ArgumentError(ValueError):
    """The reference's ArgumentError."""


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class ConvergenceError(RuntimeError):
    """An eigensolver's "did not converge" error."""


class BreakdownError(RuntimeError):
    """An eigensolver breakdown (orthogonalization failure)."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


class UnsupportedError(RuntimeError):
    """A combination this build does not serve."""


class AbortError(RuntimeError):
    """A solve stopped by its callback (the callback's own exception is re-raised
    in its place when there is one)."""


_ERR = {KFC_EARG: ArgumentError, KFC_EDIM: DimensionMismatch, KFC_ENOCONV: ConvergenceError,
        KFC_EBREAKDOWN: BreakdownError, KFC_EDEVICE: DeviceError, KFC_ENOMEM: DeviceError,
        KFC_EUNSUPPORTED: UnsupportedError, KFC_EABORT: AbortError}


class FactorDesc(C.Structure):
    _fields_ = [("axis", C.c_int32), ("kind", C.c_int32), ("d", C.c_int64), ("is_complex", C.c_int32),
                ("nbands", C.c_int32), ("offsets", C.POINTER(C.c_int64)), ("weights", C.POINTER(C.c_double)),
                ("shift", C.c_int64), ("factor_re", C.c_double), ("factor_im", C.c_double),
                ("matrix", C.POINTER(C.c_double)), ("diag", C.POINTER(C.c_double))]


class TermDesc(C.Structure):
    _fields_ = [("coeff_re", C.c_double), ("coeff_im", C.c_double), ("nfactors", C.c_int32),
                ("factors", C.POINTER(FactorDesc))]


class OpDesc(C.Structure):
    _fields_ = [("ndims", C.c_int32), ("dims", C.POINTER(C.c_int64)), ("dtype", C.c_int32),
                ("constant_re", C.c_double), ("constant_im", C.c_double), ("nterms", C.c_int32),
                ("terms", C.POINTER(TermDesc)), ("block_capacity", C.c_int64), ("hermitian", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("n_matvec", C.c_int64), ("n_matvec_f32", C.c_int64),
                ("n_apply_calls", C.c_int64), ("seconds_total", C.c_double), ("seconds_apply", C.c_double), ("seconds_apply_gpu", C.c_double),
                ("max_residual", C.c_double), ("n_converged", C.c_int32), ("warm", C.c_int32),
                ("degree_last", C.c_int32), ("retried", C.c_int32), ("seconds_ortho_gpu", C.c_double),
                ("bytes_ortho", C.c_double), ("n_ortho_calls", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class IterInfo(C.Structure):
    """kfc_iter_info: one iteration of a LOBPCG or ChebyshevFilter solve."""
    _fields_ = [("solver", C.c_int32), ("precision", C.c_int32), ("iteration", C.c_int64), ("n_matvec", C.c_int64),
                ("width", C.c_int32), ("n_locked", C.c_int32), ("active0", C.c_int32), ("degree", C.c_int32),
                ("lambda_", C.POINTER(C.c_double)), ("residuals", C.POINTER(C.c_double)),
                ("lambda_cut", C.c_double), ("lambda_max", C.c_double)]


CALLBACK = C.CFUNCTYPE(C.c_int, C.POINTER(IterInfo), C.c_void_p)

# every symbol include/kfc.h declares, with its signature (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
SIGNATURES = {
    "kfc_ctx_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "kfc_ctx_destroy": (C.c_int, [_P]),
    "kfc_last_error": (C.c_char_p, [_P]),
    "kfc_ctx_stream": (_P, [_P]),
    "kfc_ctx_set_callback": (C.c_int, [_P, CALLBACK, _P]),
    "kfc_op_create": (C.c_int, [_P, C.POINTER(OpDesc), C.POINTER(_P)]),
    "kfc_op_destroy": (C.c_int, [_P]),
    "kfc_op_size": (C.c_int64, [_P]),
    "kfc_op_fused": (C.c_int, [_P]),
    "kfc_op_stencil_get": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "kfc_op_stencil_set": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "kfc_op_fill_scratch": (C.c_int, [_P, C.c_double]),
    "kfc_op_work": (C.c_int, [_P, _D, _D]),
    "kfc_apply": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int64, C.c_int32]),
    "kfc_apply_noalloc": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int64, C.c_int32]),
    "kfc_apply_cheb_dots": (C.c_int, [_P, _P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_int32, _D,
                                      C.POINTER(C.c_int32)]),
    "kfc_sync": (C.c_int, [_P]),
    "kfc_op_diagonal": (C.c_int, [_P, _D]),
    "kfc_op_gershgorin": (C.c_int, [_P, _D]),
    "kfc_op_product_seed": (C.c_int, [_P, C.c_int64, _D]),
    "kfc_op_diagonal_device": (C.c_int, [_P, _P]),
    "kfc_complement_solve": (C.c_int, [_P, _P, _P, C.c_int64, _P, C.c_int64, _D, C.c_double, C.c_int64, C.c_double,
                                       _D, _P, C.POINTER(C.c_int64), _D, C.POINTER(C.c_int32), _D]),
    "kfc_apply_affine": (C.c_int, [_P, _P, _P, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, C.c_double]),
    "kfc_apply_interaction": (C.c_int, [_P, C.c_double, _P, C.c_double, _P, _P, C.c_int32, _P, _P, C.c_int64,
                                        C.c_int64, C.c_double, C.c_double, C.c_double, C.c_double]),
    "kfc_op_product_seed_device": (C.c_int, [_P, C.c_int64, _P, C.c_int64]),
    "kfc_lanczos": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_uint64, C.c_int32, C.c_int32, _D, _D,
                              C.POINTER(Stats)]),
    "kfc_lobpcg": (C.c_int, [_P, C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int32, C.c_double, C.c_int32, _D,
                             C.c_uint64, C.c_int32, _D, _D, C.POINTER(Stats)]),
    "kfc_lobpcg_batch": (C.c_int, [C.c_int32, C.POINTER(_P), C.c_int64, C.c_int64, C.c_double, C.c_int64, C.c_int32,
                                   C.c_double, C.c_int32, C.POINTER(_D), C.c_uint64, C.c_int32, _D, C.POINTER(_D),
                                   C.POINTER(Stats)]),
    "kfc_sandwich": (C.c_int, [_P, C.c_void_p, C.c_int64, C.c_int64, _D]),
    "kfc_lobpcg_lockstep": (C.c_int, [C.c_int32, C.POINTER(_P), C.c_int64, C.c_int64, C.c_double, C.c_int64,
                                      C.c_int32, C.c_double, C.c_int32, C.POINTER(_D), C.c_uint64, C.c_int32, _D,
                                      C.POINTER(_D), C.POINTER(Stats)]),
    "kfc_lobpcg_lockstep_subspace": (C.c_int, [C.c_int32, C.POINTER(_P), C.c_int64, C.c_int64, C.c_double, C.c_int64,
                                               C.c_int32, C.c_double, C.c_void_p, C.c_int64, C.c_int32,
                                               C.POINTER(C.POINTER(C.c_int64)), C.POINTER(_D), C.c_int32, C.c_uint64,
                                               C.c_int32, _D, C.POINTER(_D), C.POINTER(Stats)]),
    "kfc_chebyshev": (C.c_int, [_P, C.c_int64, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_int32, C.c_double,
                                C.c_int32, C.c_double, C.c_uint64, C.c_int32, _D, _D, C.POINTER(Stats)]),
    "kfc_dense": (C.c_int, [_P, C.c_int64, _D, _D, C.POINTER(Stats)]),
    "kfc_dense_batch": (C.c_int, [C.c_int32, C.POINTER(_P), C.c_int64, _D, C.POINTER(_D), C.POINTER(Stats)]),
    "kfc_residuals": (C.c_int, [_P, C.c_int64, _D, _D, _D]),
    "kfc_chebyshev_batch": (C.c_int, [C.c_int32, C.POINTER(_P), C.c_int64, C.c_int64, C.c_int32, C.c_double,
                                      C.c_int64, C.c_uint64, C.c_int32, _D, C.POINTER(_D), C.POINTER(Stats)]),
    "kfc_mlp_create": (C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32, C.c_uint64, C.POINTER(_P)]),
    "kfc_mlp_destroy": (C.c_int, [_P]),
    "kfc_mlp_set_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "kfc_mlp_get_params": (C.c_int, [_P, C.POINTER(C.c_float)]),
    "kfc_mlp_num_params": (C.c_int64, [_P]),
    "kfc_mlp_forward": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "kfc_mlp_backward": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)]),
    "kfc_mlp_train": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_int64, C.c_int64,
                                C.c_float, C.c_uint64, C.POINTER(C.c_float)]),
}

_lib = None


def lib():
    """The loaded libkfc.so with typed entry points (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libkfc.so is not built at {LIB_PATH}; run __graft_entry__.build()")
        # torch carries its own HIP runtime, which finds the GPU only when it starts
        # before this library's (the device blocks of the Python API are torch
        # tensors): start it first
        try:
            import torch
            torch.cuda.is_available()
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
        _lib = L
    return _lib


def missing_symbols():
    """Symbols declared in include/kfc.h that the loaded library does not export."""
    L = lib()
    return [n for n in SIGNATURES if not hasattr(L, n)]


def check(rc, ctx=None):
    if rc != KFC_OK:
        msg = lib().kfc_last_error(ctx).decode(errors="replace")
        raise _ERR.get(rc, RuntimeError)(msg)


class Context:
    """A device context (HIP device + stream); one per device per process."""

    def __init__(self, device=0):
        h = _P()
        check(lib().kfc_ctx_create(C.byref(h), int(device)), None)
        self.handle = h
        self.device = device
        self._torch_stream = None

    def sync(self):
        check(lib().kfc_sync(self.handle), self.handle)

    @property
    def stream(self):
        return lib().kfc_ctx_stream(self.handle)

    def torch_stream(self):
        """The library's stream as a torch stream (torch.cuda.ExternalStream)."""
        if self._torch_stream is None:
            import torch
            self._torch_stream = torch.cuda.ExternalStream(self.stream, device=torch.device("cuda", self.device))
        return self._torch_stream

    def wait_torch(self):
        """Order the library's stream after the work torch has queued on its current
        stream of this device: the library may then read tensors torch just wrote and
        write tensors torch's caching allocator just handed out."""
        import torch
        self.torch_stream().wait_stream(torch.cuda.current_stream(torch.device("cuda", self.device)))

    def torch_wait(self):
        """Order torch's current stream of this device after the work queued on the
        library's stream, so torch sees the library's results."""
        import torch
        torch.cuda.current_stream(torch.device("cuda", self.device)).wait_stream(self.torch_stream())

    def __del__(self):
        try:
            if self.handle and _lib is not None:
                _lib.kfc_ctx_destroy(self.handle)
        except Exception:
            pass


_contexts = {}


def context(device=0):
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]
