"""Parity of the HIP apply (kfc_apply through the C ABI) with the oracle's
per-term chain (src/apply.jl:370-432) on the same operators and inputs."""
import numpy as np
import pytest

from oracle import kfc_oracle as O
from tests.helpers import oracle_terms, random_ops, to_kfc, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TOL64 = 1e-13      # FP64 apply: relative to max |y|
TOL32 = 2e-6       # FP32 apply: relative to max |y|


def _run(H, X, precision="f64"):
    dt = {("f64", False): torch.float64, ("f64", True): torch.complex128,
          ("f32", False): torch.float32, ("f32", True): torch.complex64}[(precision, np.iscomplexobj(X) or H.dtype == np.complex128)]
    Xd = torch.tensor(np.ascontiguousarray(X.T), dtype=dt, device="cuda:0")
    Yd = torch.empty_like(Xd)
    H.mul(Yd, Xd, precision)
    H.sync()
    return Yd.cpu().numpy().T


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("dims", [[3, 4, 2], [5, 6, 4], [7], [9, 5], [4, 3, 2, 3]])
def test_chain_path_random(dims, cplx):
    rng = np.random.default_rng(hash((tuple(dims), cplx)) % 2**32)
    terms = random_ops(rng, dims, 6, cplx)
    H = to_kfc(dims, terms, constant=0.37)
    Ho = to_oracle(dims, terms, constant=0.37)
    X = rng.standard_normal((H.N, 5)) + (1j * rng.standard_normal((H.N, 5)) if cplx else 0)
    assert _rel(_run(H, X), O.apply_chain(Ho, X)) < TOL64
    assert _rel(_run(H, X, "f32"), O.apply_chain(Ho, X)) < 50 * TOL32


def _fusable_terms(rng, dims, cplx):
    """Kept terms whose stencil weights are rank-1 per position: constant band
    weights everywhere, one band with per-row weights along axis 0, and
    varying single-axis diagonal terms (fused into the diagonal)."""
    r = (lambda: complex(*rng.standard_normal(2))) if cplx else (lambda: float(rng.standard_normal()))
    K = len(dims)
    terms = []
    for a in range(K):
        terms.append((r(), [(a, "diag", np.array([r() for _ in range(dims[a])]))]))
        terms.append((r(), [(a, "banded", ([-1, 0, 1], [np.full(dims[a], r()) for _ in range(3)]))]))
    for a in range(K):
        for b in range(a + 1, K):
            terms.append((r(), [(a, "banded", ([-1, 1], [np.full(dims[a], r()) for _ in range(2)])),
                                (b, "shift", (int(rng.integers(-1, 2)), r()))]))
    terms.append((r(), [(0, "banded", ([1], [np.array([r() for _ in range(dims[0])])]))]))
    return terms


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("dims", [[70, 9, 11], [64, 4, 3], [130, 7], [200], [33, 17, 5]])
def test_fused_path_random(dims, cplx):
    """Banded/shift/diagonal terms with offsets in [-1, 1] run as one fused pass."""
    rng = np.random.default_rng(hash((tuple(dims), cplx, 1)) % 2**32)
    terms = _fusable_terms(rng, dims, cplx)
    H = to_kfc(dims, terms, constant=-0.5)
    assert H.fused
    Ho = to_oracle(dims, terms, constant=-0.5)
    X = rng.standard_normal((H.N, 3)) + (1j * rng.standard_normal((H.N, 3)) if cplx else 0)
    assert _rel(_run(H, X), O.apply_chain(Ho, X)) < TOL64
    assert _rel(_run(H, X, "f32"), O.apply_chain(Ho, X)) < 50 * TOL32


STENCIL_SHAPES = [(0, 64), (1, 16), (1, 64), (2, 16), (2, 64), (2, 128), (3, 16), (3, 64), (3, 128), (4, 16), (4, 64),
                  (4, 128), (5, 16), (5, 64), (5, 128)]


@pytest.mark.parametrize("variant,march", STENCIL_SHAPES)
@pytest.mark.parametrize("dims,cplx", [([70, 19, 11], False), ([64, 33, 40], False), ([33, 17, 5], True),
                                       ([130, 7], False), ([200], True), ([65, 20, 150], False)])
def test_stencil_variants(dims, cplx, variant, march):
    """Every launch shape of the fused pass on constant-weight stencils (the plane
    sweep's domain), with separable and product diagonals, ragged patch edges."""
    rng = np.random.default_rng(hash((tuple(dims), cplx, 2)) % 2**32)
    terms = _fusable_terms(rng, dims, cplx)[:-1]
    if len(dims) >= 2:
        terms.append((0.7, [(0, "diag", rng.standard_normal(dims[0])), (1, "diag", rng.standard_normal(dims[1]))]))
    H = to_kfc(dims, terms, constant=0.25)
    assert H.fused
    H.set_stencil(variant, march)
    Ho = to_oracle(dims, terms, constant=0.25)
    X = rng.standard_normal((H.N, 3)) + (1j * rng.standard_normal((H.N, 3)) if cplx else 0)
    Yo = O.apply_chain(Ho, X)
    assert _rel(_run(H, X), Yo) < TOL64
    assert _rel(_run(H, X, "f32"), Yo) < 50 * TOL32
    assert H.stencil == (variant, march)


@pytest.mark.parametrize("variant,march", [(2, 64), (4, 64), (5, 16), (5, 64), (5, 128)])
@pytest.mark.parametrize("dims", [[70, 20, 11], [64, 33, 40], [66, 17, 130]])
def test_stencil_cube_corners(dims, variant, march):
    """A kept term with offsets on all three axes reaches the eight corners of the
    stencil cube, outside the 19-position set (the plane sweeps then run on all 27
    positions); with a per-row band along axis 0 beside it, every weight constant."""
    rng = np.random.default_rng(hash((tuple(dims), variant, march, 5)) % 2**32)
    terms = _fusable_terms(rng, dims, False)[:-1]
    terms.append((0.31, [(0, "banded", ([-1, 1], [np.full(dims[0], 0.8), np.full(dims[0], -1.3)])),
                         (1, "shift", (1, 0.6)), (2, "banded", ([-1, 1], [np.full(dims[2], 1.1), np.full(dims[2], 0.4)]))]))
    H = to_kfc(dims, terms, constant=-0.125)
    assert H.fused
    H.set_stencil(variant, march)
    Ho = to_oracle(dims, terms, constant=-0.125)
    X = rng.standard_normal((H.N, 3))
    Yo = O.apply_chain(Ho, X)
    assert _rel(_run(H, X), Yo) < TOL64
    assert _rel(_run(H, X, "f32"), Yo) < 50 * TOL32
    assert H.stencil == (variant, march)


@pytest.mark.parametrize("variant,march", STENCIL_SHAPES)
@pytest.mark.parametrize("dims", [[70, 19, 11], [65, 20, 150], [130, 7], [200]])
def test_cheb_step_dots(dims, variant, march):
    """The Chebyshev step C = a H B + b B + g A with its inner products (B.B, B.C):
    C against the oracle's chain, the sums against numpy's over the same stored
    vectors, for every launch shape (the plane sweeps fuse the sums into the apply,
    the other shapes take a separate pass), FP64 and FP32 blocks, with and without a
    product diagonal."""
    rng = np.random.default_rng(hash((tuple(dims), variant, march, 3)) % 2**32)
    terms = _fusable_terms(rng, dims, False)[:-1]
    if len(dims) >= 2:
        terms.append((0.7, [(0, "diag", rng.standard_normal(dims[0])), (1, "diag", rng.standard_normal(dims[1]))]))
    H = to_kfc(dims, terms, constant=0.25)
    assert H.fused
    H.set_stencil(variant, march)
    Ho = to_oracle(dims, terms, constant=0.25)
    B, A = rng.standard_normal(H.N), rng.standard_normal(H.N)
    for a, b, g, use_a in ((0.37, -1.2, 0.0, False), (0.74, -2.4, -1.0, True)):
        Co = a * O.apply_chain(Ho, B[:, None])[:, 0] + b * B + (g * A if use_a else 0)
        for prec, tol in (("f64", TOL64), ("f32", 50 * TOL32)):
            dt = torch.float64 if prec == "f64" else torch.float32
            Bt, At = (torch.tensor(v, dtype=dt, device="cuda:0") for v in (B, A))
            Ct = torch.empty_like(Bt)
            d, fused = H.cheb_step_dots(Ct, Bt, At if use_a else None, a, b, g, prec)
            assert fused == (variant in (2, 3, 4, 5))
            Cd, Bd = Ct.double().cpu().numpy(), Bt.double().cpu().numpy()
            assert _rel(Cd, Co) < tol
            ref = np.array([Bd @ Bd, Bd @ Cd])
            assert np.all(np.abs(d - ref) <= 1e-12 * np.linalg.norm(Bd) * np.array([np.linalg.norm(Bd), np.linalg.norm(Cd)]))


@pytest.mark.parametrize("variant,march", [(-1, 0)] + STENCIL_SHAPES)
@pytest.mark.parametrize("N", [16, 33, 64, 100])
def test_three_fluxonia_fused_parity(N, variant, march):
    from kfc import configs
    th = configs.three_fluxonia_theta(11)
    H = configs.three_fluxonia_operator(th, N=N)
    assert H.fused
    H.set_stencil(variant, march)
    Ho = O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=N, order=2)
    rng = np.random.default_rng(N)
    X = rng.standard_normal((H.N, 2))
    Yo = O.apply_chain(Ho, X)
    assert _rel(_run(H, X), Yo) < TOL64
    assert _rel(_run(H, X, "f32"), Yo) < TOL32 * 10


def test_leading_dimension_and_views():
    from kfc import configs
    th = configs.three_fluxonia_theta(3)
    H = configs.three_fluxonia_operator(th, N=24)
    Ho = O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=24, order=2)
    rng = np.random.default_rng(1)
    big = torch.tensor(rng.standard_normal((4, H.N + 37)), device="cuda:0")
    X = big[:, 5:5 + H.N]
    Y = torch.zeros_like(big)[:, 3:3 + H.N]
    H.mul(Y, X)
    H.sync()
    Yo = O.apply_chain(Ho, X.cpu().numpy().T).T
    assert _rel(Y.cpu().numpy(), Yo) < TOL64


def test_alias_rejected():
    import kfc
    from kfc import configs
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(3), N=12)
    X = torch.zeros((2, H.N), dtype=torch.float64, device="cuda:0")
    with pytest.raises(kfc.ArgumentError):
        H.mul(X, X)


def test_matrix_equals_kron_reference():
    rng = np.random.default_rng(5)
    dims = [4, 3, 5]
    terms = random_ops(rng, dims, 5, True)
    H = to_kfc(dims, terms, 1.5)
    Ho = to_oracle(dims, terms, 1.5)
    M = _run(H, np.eye(H.N, dtype=complex))
    # the constant is one of the fused terms (a term with no factors)
    assert np.allclose(M, O.kron_reference(dims, Ho.terms + Ho.diagterms), atol=1e-12)
    assert np.allclose(M, O.kron_reference(dims, oracle_terms(dims, terms), 1.5), atol=1e-12)


def test_derived_quantities():
    from kfc import configs
    th = configs.three_fluxonia_theta(3)
    H = configs.three_fluxonia_operator(th, N=10)
    Ho = O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=10, order=2)
    assert abs(H.gershgorin_bound() - O.gershgorin_bound(Ho)) < 1e-9 * abs(O.gershgorin_bound(Ho))
    assert np.allclose(H.diagonal(), O.full_diagonal(Ho), atol=1e-10)
    P, Po = H.product_seed(6), O.product_seed(Ho, 6)
    # same product states up to each column's sign
    assert np.allclose(np.abs(np.sum(P * Po, axis=0)), 1.0, atol=1e-10)


@pytest.mark.parametrize("dims", [[40, 40, 30], [7, 9, 5, 3], [16, 33, 2], [64, 3, 17]])
def test_dense_pair_mfma_path(dims):
    """Two-factor dense terms (a branch node's kept-basis ports) on MFMA tiles."""
    import kfc
    rng = np.random.default_rng(sum(dims))
    K = len(dims)
    terms = []
    pairs = [(0, 1), (1, 2), (0, K - 1)]
    for t in range(6):
        p, q = pairs[t % len(pairs)]
        A = rng.standard_normal((dims[p], dims[p]))
        B = rng.standard_normal((dims[q], dims[q]))
        if t >= 3:   # share a factor with an earlier term: merged on the host
            A = terms[t - 3][1][0][2]
        terms.append((float(rng.standard_normal()), [(p, "dense", A), (q, "dense", B)]))
    terms.append((0.3, [(0, "diag", rng.standard_normal(dims[0]))]))
    H = to_kfc(dims, terms, constant=0.25)
    assert not H.fused
    Ho = to_oracle(dims, terms, constant=0.25)
    X = rng.standard_normal((H.N, 4))
    Yo = O.apply_chain(Ho, X)
    assert _rel(_run(H, X), Yo) < TOL64
    assert _rel(_run(H, X, "f32"), Yo) < 20 * TOL32


def test_zero_pi_parent_operator_parity():
    """The zero-pi tree's parent [40, 40, 30] (dense ports) against the oracle chain."""
    import kfc
    from kfc import configs
    from tests.helpers import oracle_from_kfc_terms
    th = configs.zero_pi_theta(3)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    root = configs.zero_pi_tree(E_Cz)
    H, kids = kfc.operator(root, spec)
    assert H.dims == [40, 40, 30] and not H.fused
    Ho = oracle_from_kfc_terms(H)
    rng = np.random.default_rng(0)
    X = rng.standard_normal((H.N, 20))
    assert _rel(_run(H, X), O.apply_chain(Ho, X)) < TOL64


def test_three_fluxonia_full_size_parity():
    """The bench node at its full size, 512^3 (134 M states): the fused stencil apply
    in FP64 against the C restatement of the reference's per-term chain
    (oracle/apply_chain.c) on the same random vector, the FP32 apply within its
    tolerance, and the symmetry <u, H v> = <H u, v> of the real Hermitian operator."""
    from kfc import configs
    from oracle import c_apply
    if not c_apply.available():
        pytest.skip("the oracle's C restatement is not built (run __graft_entry__.build())")
    th = configs.three_fluxonia_theta(20261203)
    N = 512
    H = configs.three_fluxonia_operator(th, N=N)
    assert H.fused and H.N == N ** 3
    rng = np.random.default_rng(17)
    x = rng.standard_normal(H.N)
    xd = torch.from_numpy(x).to("cuda:0")
    y = (H @ xd)
    H.sync()
    y = y.cpu().numpy()
    Ho = O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=N, order=2)
    yref = c_apply.CApply(Ho)(x)
    del Ho
    assert _rel(y, yref) < TOL64
    y32 = H @ xd.float()
    H.sync()
    assert _rel(y32.cpu().numpy().astype(np.float64), yref) < 50 * TOL32
    del y32, y, yref
    u = torch.from_numpy(rng.standard_normal(H.N)).to("cuda:0")
    Hu = H @ u
    Hx = H @ xd
    H.sync()
    a, b = torch.dot(u, Hx).item(), torch.dot(Hu, xd).item()
    assert abs(a - b) <= 1e-10 * (torch.linalg.norm(u) * torch.linalg.norm(Hx)).item()
