"""The warm-start network over kfc_mlp_* (BF16-MFMA forward/backward kernels,
FP32 accumulation, Adam on the device).  No reference counterpart: its output
feeds the X0 argument of LOBPCG (src/lobpcg.jl:217, src/materialize.jl:670)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

F = C.POINTER(C.c_float)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class MLP:
    """GELU MLP with layer widths ``dims`` (e.g. 6 -> 256 -> 256 -> 1024)."""

    def __init__(self, dims=(6, 256, 256, 1024), seed=0, device=0):
        self.ctx = L.context(device)
        self.dims = [int(d) for d in dims]
        arr = (C.c_int32 * len(self.dims))(*self.dims)
        h = C.c_void_p()
        L.check(L.lib().kfc_mlp_create(self.ctx.handle, arr, len(self.dims), int(seed), C.byref(h)), self.ctx.handle)
        self.handle = h
        self.nparams = int(L.lib().kfc_mlp_num_params(h))

    def __del__(self):
        try:
            if getattr(self, "handle", None) and L._lib is not None:
                L._lib.kfc_mlp_destroy(self.handle)
        except Exception:
            pass

    def shapes(self):
        out = []
        for i in range(1, len(self.dims)):
            out += [(self.dims[i], self.dims[i - 1]), (self.dims[i],)]
        return out

    def get_params(self):
        p = np.empty(self.nparams, dtype=np.float32)
        L.check(L.lib().kfc_mlp_get_params(self.handle, p.ctypes.data_as(F)), self.ctx.handle)
        out, o = [], 0
        for s in self.shapes():
            n = int(np.prod(s))
            out.append(p[o:o + n].reshape(s).copy())
            o += n
        return out

    def set_params(self, params):
        p = _f32(np.concatenate([np.asarray(x, dtype=np.float32).reshape(-1) for x in params]))
        L.check(L.lib().kfc_mlp_set_params(self.handle, p.ctypes.data_as(F)), self.ctx.handle)

    def forward(self, X):
        X = _f32(X)
        Y = np.empty((X.shape[0], self.dims[-1]), dtype=np.float32)
        L.check(L.lib().kfc_mlp_forward(self.handle, X.ctypes.data_as(F), X.shape[0], Y.ctypes.data_as(F)),
                self.ctx.handle)
        return Y

    def backward(self, X, dY):
        X, dY = _f32(X), _f32(dY)
        g = np.empty(self.nparams, dtype=np.float32)
        L.check(L.lib().kfc_mlp_backward(self.handle, X.ctypes.data_as(F), dY.ctypes.data_as(F), X.shape[0],
                                         g.ctypes.data_as(F)), self.ctx.handle)
        out, o = [], 0
        for s in self.shapes():
            n = int(np.prod(s))
            out.append(g[o:o + n].reshape(s).copy())
            o += n
        return out

    def train(self, X, T, batch=256, steps=1000, lr=1e-3, seed=0):
        X, T = _f32(X), _f32(T)
        loss = C.c_float()
        L.check(L.lib().kfc_mlp_train(self.handle, X.ctypes.data_as(F), T.ctypes.data_as(F), X.shape[0], int(batch),
                                      int(steps), float(lr), int(seed), C.byref(loss)), self.ctx.handle)
        return loss.value
