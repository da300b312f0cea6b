"""Derived quantities formed on the device: the full diagonal (kfc_op_diagonal_device,
src/apply.jl:455-478), the fused diagonal's range inside the Gershgorin bound
(kfc_op_gershgorin, src/apply.jl:586-593) and the product seed
(kfc_op_product_seed_device, src/apply.jl:655-685).  Each equals the host form bit for
bit; the host forms are pinned against the oracle here and in tests/test_gpu_apply.py.
Cases: real and complex operators, diagonals that are not sums over axes, ties in the
Kronecker sum (two identical axes) and in the diagonal (the radix-select branch)."""
import ctypes as C

import numpy as np
import pytest

from helpers import random_ops, to_kfc, to_oracle
from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu


def _seed_dev(H, k):
    return H.product_seed_device(k).cpu().numpy()


def _nonseparable_terms(rng, dims, cplx):
    terms = random_ops(rng, dims, 8, cplx)
    r = (lambda n: rng.standard_normal(n) + 1j * rng.standard_normal(n)) if cplx else rng.standard_normal
    # diagonal products on two and three axes: the fused diagonal is not a sum over axes
    terms.append((0.7, [(0, "diag", r(dims[0])), (2, "diag", r(dims[2]))]))
    terms.append((-0.4, [(0, "diag", r(dims[0])), (1, "diag", r(dims[1])), (2, "diag", r(dims[2]))]))
    return terms


@pytest.mark.parametrize("cplx", [False, True])
def test_diagonal_device_equals_host(cplx):
    rng = np.random.default_rng(11 + cplx)
    dims = [7, 5, 6]
    terms = _nonseparable_terms(rng, dims, cplx)
    H = to_kfc(dims, terms, constant=0.3)
    D = H.diagonal_device().cpu().numpy()
    assert np.array_equal(D, H.diagonal())
    Do = np.real(O.full_diagonal(to_oracle(dims, terms, 0.3)))
    assert np.max(np.abs(D - Do)) <= 1e-12 * max(1.0, np.max(np.abs(Do)))


@pytest.mark.parametrize("cplx", [False, True])
def test_gershgorin_nonseparable_diagonal(cplx):
    """The fused diagonal's max/min real part and max |imaginary part| reduced on the
    device (complex diagonals carry imaginary parts here)."""
    rng = np.random.default_rng(21 + cplx)
    dims = [9, 8, 7]
    terms = _nonseparable_terms(rng, dims, cplx)
    H = to_kfc(dims, terms, constant=-1.25)
    g = O.gershgorin_bound(to_oracle(dims, terms, -1.25))
    assert abs(H.gershgorin_bound() - g) <= 1e-12 * abs(g)


@pytest.mark.parametrize("k", [1, 6, 20, 64])
def test_product_seed_three_fluxonia(k):
    from kfc import configs
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(3), N=10)
    P = _seed_dev(H, k)
    assert np.array_equal(P, H.product_seed(k))
    assert np.allclose(P.T @ P, np.eye(k), atol=1e-12)


def test_product_seed_ties_in_the_kronecker_sum():
    """Two identical axes: every sum off the diagonal of the (i, j) grid appears twice,
    so the best-first walk must order ties by flat index exactly as the scan does."""
    rng = np.random.default_rng(5)
    A = rng.standard_normal((8, 8))
    A = A + A.T
    B = rng.standard_normal((3, 3))
    B = B + B.T
    Cc = rng.standard_normal((8, 8))
    dims = [8, 8, 3]
    terms = [(1.0, [(0, "dense", A)]), (1.0, [(1, "dense", A)]), (0.5, [(2, "dense", B)]),
             (0.1, [(0, "dense", Cc + Cc.T), (1, "dense", Cc + Cc.T)])]
    H = to_kfc(dims, terms)
    Ho = to_oracle(dims, terms)
    for k in (1, 5, 17, 40, 192):
        P = _seed_dev(H, k)
        assert np.array_equal(P, H.product_seed(k))
        assert np.allclose(np.abs(np.sum(P * O.product_seed(Ho, k), axis=0)), 1.0, atol=1e-10)


def test_product_seed_complex_tree_top():
    """A coupled-transmon top on its children's eigenbases (complex; every axis an
    eigenbasis, so the unit vectors of the smallest diagonal entries)."""
    from kfc import transmon_sweep as T
    from kfc.tree import operator
    inst = T.draw(1, seed=7)[0]
    inst["N1"], inst["N2"] = 256, 256
    H, _ = operator(T.tree_of(inst, k_leaf=8, k_top=20), T.spec_of(inst))
    assert H.dtype == np.complex128
    for k in (1, 9, 30, 64):
        P = _seed_dev(H, k)
        assert np.array_equal(P, H.product_seed(k))


def _unit_seed(D, k):
    order = np.lexsort((np.arange(D.size), D))[:k]
    P = np.zeros((D.size, k))
    P[order, np.arange(k)] = 1.0
    return P


def test_product_seed_radix_select_ties():
    """Every axis an eigenbasis and a diagonal that repeats each value 48 times: the k
    smallest entries, ties by index (the order partialsortperm gives)."""
    rng = np.random.default_rng(9)
    dims = [16, 12]
    d0 = np.repeat(np.round(rng.standard_normal(4), 3), 4)
    C0 = rng.standard_normal((16, 16))
    np.fill_diagonal(C0, 0.0)
    C1 = rng.standard_normal((12, 12))
    np.fill_diagonal(C1, 0.0)
    terms = [(1.0, [(0, "diag", d0)]), (0.2, [(0, "dense", C0 + C0.T), (1, "dense", C1 + C1.T)])]
    H = to_kfc(dims, terms)
    D = H.diagonal()
    for k in (1, 7, 33, 100, 192):
        P = _seed_dev(H, k)
        assert np.array_equal(P, H.product_seed(k))
        assert np.array_equal(P, _unit_seed(D, k))


def test_product_seed_radix_select_large():
    """2^18 distinct diagonal entries (a three-axis diagonal product), 200 selected."""
    rng = np.random.default_rng(13)
    dims = [64, 64, 64]
    C0 = rng.standard_normal((64, 64))
    np.fill_diagonal(C0, 0.0)
    terms = [(1.0, [(0, "diag", rng.uniform(1, 2, 64)), (1, "diag", rng.uniform(1, 2, 64)),
                    (2, "diag", rng.uniform(1, 2, 64))]),
             (0.1, [(0, "dense", C0 + C0.T), (2, "dense", C0 + C0.T)])]
    H = to_kfc(dims, terms)
    D = H.diagonal()
    P = _seed_dev(H, 200)
    assert np.array_equal(P, _unit_seed(D, 200))


def test_seed_device_rejects():
    import kfc
    import torch
    from kfc import _lib as L
    from kfc import configs
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(1), N=8)
    X = torch.zeros(4, H.N, dtype=torch.float64, device="cuda")
    for k, ld in ((0, H.N), (H.N + 1, H.N), (4, H.N - 1)):
        with pytest.raises(kfc.ArgumentError):
            L.check(L.lib().kfc_op_product_seed_device(H.handle, k, C.c_void_p(X.data_ptr()), ld), H.ctx.handle)
