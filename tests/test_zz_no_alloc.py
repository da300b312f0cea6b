"""kfc_apply_noalloc, the variant of kfc_apply for callers that cannot allocate device
memory (no counterpart in the reference), called through the C ABI: on every path
(chain with its scratch blocks, fused stencil, dense kept-basis pairs) and at both
precisions it gives kfc_apply's result bit for bit, on two calls in a row with the
second smaller; it allocates no device memory where kfc_apply would; oversized blocks,
complex blocks on a real operator and absent FP32 tables are refused."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import random_ops, to_kfc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _call(fn, H, X, Y, prec):
    from kfc import _lib as L
    b, ld = (1, H.N) if X.dim() == 1 else (X.shape[0], X.stride(0))
    H.ctx.wait_torch()
    rc = fn(H.handle, C.c_void_p(X.data_ptr()), C.c_void_p(Y.data_ptr()), b, ld, prec)
    L.check(rc, H.ctx.handle)
    H.ctx.torch_wait()
    return Y


def _both(H, X, prec):
    from kfc import _lib as L
    Ya = _call(L.lib().kfc_apply, H, X, torch.empty_like(X), prec)
    Yn = _call(L.lib().kfc_apply_noalloc, H, X, torch.full_like(X, float("nan")), prec)
    torch.cuda.synchronize()
    return Ya.cpu().numpy(), Yn.cpu().numpy()


def _ops():
    import kfc
    from kfc import configs
    rng = np.random.default_rng(7)
    dims = [9, 8, 7]
    chain_c = to_kfc(dims, random_ops(rng, dims, 5, True), constant=0.4)
    chain_r = to_kfc(dims, random_ops(rng, dims, 5, False), constant=-0.2)
    fused = configs.three_fluxonia_operator(configs.three_fluxonia_theta(2), N=30)
    mk = lambda n: (lambda A: (A + A.T) / 2)(rng.standard_normal((n, n)))
    pair = kfc.KronFreeOperator([12, 10, 8], [kfc.AxisTerm(1.0, [(0, kfc.DiagonalOp(rng.uniform(0, 3, 12)))]),
                                              kfc.AxisTerm(0.3, [(0, kfc.DenseOp(mk(12))), (1, kfc.DenseOp(mk(10)))]),
                                              kfc.AxisTerm(-0.2, [(1, kfc.DenseOp(mk(10))), (2, kfc.DenseOp(mk(8)))])])
    assert fused.fused
    return {"chain_c": chain_c, "chain_r": chain_r, "fused": fused, "pair": pair}


def _block(H, b, dt, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if dt == torch.complex128:
        return torch.complex(torch.randn(b, H.N, dtype=torch.float64, device="cuda", generator=g),
                             torch.randn(b, H.N, dtype=torch.float64, device="cuda", generator=g))
    if dt == torch.complex64:
        return torch.complex(torch.randn(b, H.N, device="cuda", generator=g),
                             torch.randn(b, H.N, device="cuda", generator=g))
    return torch.randn(b, H.N, dtype=dt, device="cuda", generator=g)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_equals_apply_two_calls(prec):
    from kfc import _lib as L
    ops = _ops()
    p = L.KFC_PREC_F64 if prec == "f64" else L.KFC_PREC_F32
    for name, H in ops.items():
        cplx = H.dtype == np.complex128
        dt = {("f64", False): torch.float64, ("f32", False): torch.float32,
              ("f64", True): torch.complex128, ("f32", True): torch.complex64}[(prec, cplx)]
        # two calls in a row, the second on a smaller block (and a vector)
        for b, seed in ((6, 1), (2, 2)):
            Ya, Yn = _both(H, _block(H, b, dt, seed), p)
            assert np.array_equal(Ya, Yn), (name, prec, b)
            assert np.all(np.isfinite(Yn))
        x = _block(H, 1, dt, 3)[0]
        Ya, Yn = _both(H, x, p)
        assert np.array_equal(Ya, Yn), (name, prec, "vector")
    # the second operator smaller than the first, through the same shared array
    big, small = ops["chain_c"], to_kfc([3, 4, 2], random_ops(np.random.default_rng(9), [3, 4, 2], 4, True))
    for H, b in ((big, 5), (small, 3)):
        dt = torch.complex128 if prec == "f64" else torch.complex64
        Ya, Yn = _both(H, _block(H, b, dt, 4), p)
        assert np.array_equal(Ya, Yn)


def test_noalloc_allocates_nothing():
    """Free device memory (hipMemGetInfo) is unchanged by kfc_apply_noalloc on a chain
    operator whose kfc_apply grows two scratch blocks of 4 MiB each."""
    from kfc import _lib as L
    rng = np.random.default_rng(11)
    dims = [40, 40, 40]
    terms = random_ops(rng, dims, 3, False, kinds=("banded", "dense", "shift"))
    assert any(len(fs) > 1 for _, fs in terms)
    Ha, Hn = to_kfc(dims, terms), to_kfc(dims, terms)
    X = _block(Ha, 8, torch.float64, 5)          # N x 8 doubles = 4 MiB per scratch block
    Ya, Yn = torch.empty_like(X), torch.empty_like(X)
    Hf = _ops()["fused"]
    xf = _block(Hf, 1, torch.float64, 6)
    _call(L.lib().kfc_apply_noalloc, Hf, xf, torch.empty_like(xf), L.KFC_PREC_F64)      # the module is loaded
    torch.cuda.synchronize()
    f0, _ = torch.cuda.mem_get_info()
    _call(L.lib().kfc_apply_noalloc, Hn, X, Yn, L.KFC_PREC_F64)
    torch.cuda.synchronize()
    f1, _ = torch.cuda.mem_get_info()
    _call(L.lib().kfc_apply, Ha, X, Ya, L.KFC_PREC_F64)
    torch.cuda.synchronize()
    f2, _ = torch.cuda.mem_get_info()
    assert f0 - f1 == 0, f0 - f1
    assert f1 - f2 >= 8 << 20, f1 - f2          # the control: kfc_apply grew its two blocks
    assert np.array_equal(Ya.cpu().numpy(), Yn.cpu().numpy())


def test_noalloc_refusals():
    import kfc
    from kfc import _lib as L
    rng = np.random.default_rng(12)
    dims = [64, 64, 64]
    H = to_kfc(dims, random_ops(rng, dims, 2, True, kinds=("banded", "shift")))
    assert any(len(t.factors) > 1 for t in H.terms)
    X = _block(H, 9, torch.complex128, 7)        # 2 x 9 x 262144 x 16 B = 72 MiB of scratch
    with pytest.raises(kfc.ArgumentError, match="scratch"):
        _call(L.lib().kfc_apply_noalloc, H, X, torch.empty_like(X), L.KFC_PREC_F64)
    Y8 = _call(L.lib().kfc_apply_noalloc, H, X[:7], torch.empty_like(X[:7]), L.KFC_PREC_F64)   # 56 MiB fits
    Ya = _call(L.lib().kfc_apply, H, X[:7], torch.empty_like(X[:7]), L.KFC_PREC_F64)
    assert np.array_equal(Y8.cpu().numpy(), Ya.cpu().numpy())
    ops = _ops()
    Hr = ops["fused"]
    z = _block(Hr, 2, torch.complex128, 8)
    with pytest.raises(kfc.UnsupportedError):
        _call(L.lib().kfc_apply_noalloc, Hr, z, torch.empty_like(z), L.KFC_PREC_Z64)
    x32 = _block(ops["chain_r"], 2, torch.float32, 9)
    with pytest.raises(kfc.ArgumentError, match="FP32 tables"):
        _call(L.lib().kfc_apply_noalloc, ops["chain_r"], x32, torch.empty_like(x32), L.KFC_PREC_F32)
