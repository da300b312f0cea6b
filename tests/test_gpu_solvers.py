"""Device solvers against the oracle's Dense (numpy eigh) and the reference's
golden energies (test/fixtures/scqubits_reference.json)."""
import json
import os

import numpy as np
import pytest

from oracle import kfc_oracle as O
from tests.helpers import oracle_from_kfc_terms

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _ref(golden_dir, name):
    with open(os.path.join(golden_dir, "scqubits_reference.json")) as f:
        e = json.load(f)[name]
    return np.array(e["energies"]), e["params"]


def test_dense_two_transmons_vs_scqubits(golden_dir):
    """test/test_materialize.jl:142-160 through the device Dense solver."""
    import kfc
    from kfc import configs
    E, p = _ref(golden_dir, "two_transmons")
    sys_ = kfc.materialize(configs.two_transmons_tree(), configs.two_transmons_spec(**configs.TWO_TRANSMONS_THETA))
    En = sys_.energies()
    assert np.max(np.abs(En[:3] - E[:3])) < 1e-8
    assert np.max(np.abs(En - E)) < 5e-4


def test_transmon_leaf_vs_scqubits(golden_dir):
    import kfc
    E, p = _ref(golden_dir, "transmon")
    spec = kfc.NumericSpec(["varphi_1"], [[p["E_C"]]], [[0.0]], [0.0], [p["n_g"]], [(p["E_J"], [1], 0.0)])
    s = kfc.materialize(kfc.Node([("varphi_1", kfc.ChargeBasis(30))], keep=5), spec)
    assert np.max(np.abs(s.energies() - E)) < 1e-10


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_lanczos_small_fluxonia_vs_dense(mode):
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(2)
    H = configs.three_fluxonia_operator(th, N=14)
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))
    E, V, st = kfc.solve_stats(kfc.Lanczos(tol=1e-10, mode=mode), H, 10)
    assert np.max(np.abs(E - w[:10])) < 1e-9 * max(1.0, np.max(np.abs(w[:10])))
    r = O.true_residuals(Ho, E, V)
    assert np.max(r) < 1e-8
    assert np.allclose(V.T @ V, np.eye(10), atol=1e-10)
    assert st["max_residual"] < 1e-8


@pytest.mark.parametrize("mode", ["f64", "mixed"])
@pytest.mark.parametrize("precond", ["ritz", "jacobi", "none"])
def test_lobpcg_zero_pi_tree_small(mode, precond):
    """A small zero-pi tree: leaves Dense, parent LOBPCG vs the parent's dense spectrum."""
    import kfc
    from kfc import configs
    th = configs.zero_pi_theta(1)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    root = configs.zero_pi_tree(E_Cz, ncut=12, Nphi=64, Nzeta=10, keep=10, m=6,
                                solver=kfc.LOBPCG(tol=1e-8, preconditioner=precond, mode=mode))
    s = kfc.materialize(root, spec)
    H = s.root.operator
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))
    assert np.max(np.abs(s.energies() - w[:6])) < 1e-9 * max(1.0, np.max(np.abs(w[:6])))
    r = O.true_residuals(Ho, s.energies(), s.root.vectors)
    assert np.max(r) < 1e-8


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_lobpcg_wide_block_vs_dense(mode):
    """m = 24 (block [X R P] of about 3 (m + guard) > 64 columns): its Grams run as
    64 x 64 tiles on the matrix cores, the last tiles ragged (FP32 blocks in the mixed
    mode's first phase).  Eigenvalues and true residuals against the dense spectrum of
    the oracle's operator."""
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(3)
    H = configs.three_fluxonia_operator(th, N=12)
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))
    m = 24
    assert 3 * (m + kfc.solvers.guard_count(kfc.LOBPCG(), H.N, m)) > 64
    E, V, st = kfc.solve_stats(kfc.LOBPCG(tol=1e-8, mode=mode), H, m)
    assert np.max(np.abs(E - w[:m])) < 1e-9 * max(1.0, np.max(np.abs(w[:m])))
    assert np.max(O.true_residuals(Ho, E, V)) < 1e-7
    assert np.allclose(V.T @ V, np.eye(m), atol=1e-9)


def test_lobpcg_warm_start_and_errors():
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(5)
    H = configs.three_fluxonia_operator(th, N=10)
    E, V = kfc.solve(kfc.LOBPCG(tol=1e-8), H, 4)
    E2, V2, st = kfc.solve_stats(kfc.LOBPCG(tol=1e-8), H, 4, X0=V)
    assert st["warm"] == 1 and np.allclose(E, E2, atol=1e-8)
    with pytest.raises(kfc.ArgumentError):
        kfc.solve(kfc.LOBPCG(), H, 400)                    # N < 3m + 5
    bad = V.copy()
    bad[:, 1] = 0
    with pytest.raises(kfc.ArgumentError):
        kfc.solve(kfc.LOBPCG(), H, 4, X0=bad)              # zero column
    bad[0, 0] = np.nan
    with pytest.raises(kfc.ArgumentError):
        kfc.solve(kfc.LOBPCG(), H, 4, X0=bad)              # non-finite
    with pytest.raises(kfc.ConvergenceError):
        kfc.solve(kfc.LOBPCG(maxiter=1, tol=1e-12), H, 4)  # nonconvergence


def test_dense_complex_periodic_vs_closed_form():
    """Free rotor on a PeriodicGrid (complex node) through the device Dense solver."""
    import math
    import kfc
    N, EC, ng, order = 32, 1.1, 0.27, 8
    b = kfc.PeriodicGrid(N, order)
    H = kfc.KronFreeOperator([N], [kfc.AxisTerm(4 * EC, [(0, kfc.localop(b, ("n2", 0)))]),
                                   kfc.AxisTerm(-8 * EC * ng, [(0, kfc.localop(b, ("n", 0)))])],
                             constant=4 * EC * ng ** 2)
    E, V = kfc.solve(kfc.Dense(), H, 0)
    h = order // 2
    w1, w2 = kfc.fd_weights(order, 1), kfc.fd_weights(order, 2)
    D = 2 * math.pi / N
    ks = np.arange(-N // 2, N // 2)
    s1 = np.array([sum(w1[j + h] * math.sin(j * k * D) for j in range(-h, h + 1)) / D for k in ks])
    s2 = np.array([-sum(w2[j + h] * math.cos(j * k * D) for j in range(-h, h + 1)) / D ** 2 for k in ks])
    assert np.allclose(E, np.sort(4 * EC * (s2 - 2 * ng * s1 + ng ** 2)), atol=1e-9)


@pytest.mark.parametrize("ortho", ["svqb", "cholqr"])
@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_chebyshev_fluxonium_leaf_vs_dense(ortho, mode):
    """ChebyshevFilter on a grid leaf (the device leaf solver, src/materialize.jl:557)."""
    import kfc
    b = kfc.PhaseGrid(-10 * np.pi, 10 * np.pi, 801, 8)
    spec = kfc.NumericSpec(["varphi_1"], [[2.5]], [[0.5]], [0.5 * np.pi], [0.0], [(8.9, [1], 0.0)], 0.5 * np.pi ** 2 / 2)
    node = kfc.Node([("varphi_1", b)], keep=6, solver=kfc.ChebyshevFilter(ortho=ortho, mode=mode, extra=20))
    s = kfc.materialize(node, spec)
    H = s.root.operator
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))[:6]
    assert np.max(np.abs(s.energies() - w)) < 1e-9 * max(1, np.max(np.abs(w)))
    assert s.root.stats["max_residual"] <= max(1e-8, 10 * 2.2e-16 * H.gershgorin_bound())


def test_chebyshev_complex_periodic_transmon():
    import kfc
    N, EJ, EC, ng = 128, 30.0, 1.2, 0.25
    b = kfc.PeriodicGrid(N, 8)
    H = kfc.KronFreeOperator([N], [kfc.AxisTerm(4 * EC, [(0, kfc.localop(b, ("n2", 0)))]),
                                   kfc.AxisTerm(-8 * EC * ng, [(0, kfc.localop(b, ("n", 0)))]),
                                   kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", 1)))]),
                                   kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", -1)))])],
                             constant=4 * EC * ng ** 2)
    assert H.dtype == np.complex128
    E, V = kfc.solve(kfc.ChebyshevFilter(extra=10), H, 5)
    Ed, _ = kfc.solve(kfc.Dense(), H, 5)
    assert np.allclose(E, Ed, atol=1e-9)


@pytest.mark.parametrize("mode", ["f64", "mixed"])
@pytest.mark.parametrize("krylovdim", [0, 48])
def test_lanczos_complex_periodic_transmon(mode, krylovdim):
    """Complex Hermitian Lanczos (the general block path) against Dense, plus the
    Krylov-dimension keyword."""
    import kfc
    N, EJ, EC, ng = 300, 25.0, 1.0, 0.31
    b = kfc.PeriodicGrid(N, 8)
    H = kfc.KronFreeOperator([N], [kfc.AxisTerm(4 * EC, [(0, kfc.localop(b, ("n2", 0)))]),
                                   kfc.AxisTerm(-8 * EC * ng, [(0, kfc.localop(b, ("n", 0)))]),
                                   kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", 1)))]),
                                   kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", -1)))])],
                             constant=4 * EC * ng ** 2)
    assert H.dtype == np.complex128
    E, V, st = kfc.solve_stats(kfc.Lanczos(tol=1e-9, krylovdim=krylovdim, mode=mode), H, 6)
    Ed, _ = kfc.solve(kfc.Dense(), H, 6)
    assert np.max(np.abs(E - Ed)) < 1e-9 * max(1.0, np.max(np.abs(Ed)))
    assert np.allclose(V.conj().T @ V, np.eye(6), atol=1e-9)
    assert st["max_residual"] < 1e-8


@pytest.mark.parametrize("N", [14, 18])
def test_lanczos_real_block_paths_vs_dense(N):
    """Real Lanczos at several Krylov dimensions: short and long Chebyshev blocks,
    thick restarts and the explicit check all agree with Dense."""
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(5)
    H = configs.three_fluxonia_operator(th, N=N)
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))
    for kd in (0, 40, 64):
        E, V, st = kfc.solve_stats(kfc.Lanczos(tol=1e-10, krylovdim=kd), H, 8)
        r = O.true_residuals(Ho, E, V)
        assert np.max(r) < 1e-8
        assert np.max(np.abs(E - w[:8])) < 1e-9 * max(1.0, np.max(np.abs(w[:8])))
        assert np.allclose(V.T @ V, np.eye(8), atol=1e-10)


@pytest.mark.parametrize("mode", ["f64", "mixed"])
@pytest.mark.parametrize("variant,march", [(1, 32), (2, 32), (3, 32)])
def test_lanczos_fused_moments_vs_dense(mode, variant, march):
    """The s-step Lanczos takes each block's moments q^T T_j q from the plane
    sweep's fused sums (variants 2 and 3) or from a Gram pass (variant 1): both
    reach Dense's spectrum with the same orthonormality."""
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(7)
    H = configs.three_fluxonia_operator(th, N=18)
    H.set_stencil(variant, march)
    Ho = oracle_from_kfc_terms(H)
    w = np.linalg.eigvalsh(O.to_matrix(Ho))
    E, V, st = kfc.solve_stats(kfc.Lanczos(tol=1e-10, krylovdim=48, mode=mode), H, 8)
    assert np.max(np.abs(E - w[:8])) < 1e-9 * max(1.0, np.max(np.abs(w[:8])))
    assert np.max(O.true_residuals(Ho, E, V)) < 1e-8
    assert np.allclose(V.T @ V, np.eye(8), atol=1e-10)


def _hermitian_operator(rng, dims, cplx):
    """Every LocalOp kind in Hermitian combinations: dense Hermitian factors on two
    axes, a symmetric banded factor, a shift plus its adjoint, a real diagonal, a
    constant."""
    import kfc

    def herm(n):
        A = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
        return (A + A.conj().T) / 2
    d0, d1 = dims[0], dims[1]
    w = rng.standard_normal(d1)
    lo, hi = np.roll(w, 1), w.copy()          # y[i] = lo[i] x[i-1] + c[i] x[i] + hi[i] x[i+1], hi[i] = lo[i+1]
    f = complex(*rng.standard_normal(2)) if cplx else float(rng.standard_normal())
    terms = [kfc.AxisTerm(0.7, [(0, kfc.DenseOp(herm(d0))), (1, kfc.DenseOp(herm(d1)))]),
             kfc.AxisTerm(0.3, [(1, kfc.BandedOp(d1, [-1, 0, 1], [lo, rng.standard_normal(d1), hi]))]),
             kfc.AxisTerm(0.5, [(0, kfc.ShiftOp(d0, 1, f))]),
             kfc.AxisTerm(0.5, [(0, kfc.ShiftOp(d0, -1, np.conj(f)))]),
             kfc.AxisTerm(1.0, [(len(dims) - 1, kfc.DiagonalOp(rng.uniform(-1, 1, dims[-1])))])]
    return kfc.KronFreeOperator(dims, terms, constant=0.4)


def test_dense_host_matrix_and_batch():
    """Dense of small operators forms Matrix(H) on the host from the terms (every
    LocalOp kind, real and complex, a constant and a fused diagonal); the result
    matches the oracle's dense matrix, and kfc_dense_batch over several operators (one
    above the host limit, through the device path) matches kfc_dense one by one."""
    import kfc
    rng = np.random.default_rng(77)
    Hs = [_hermitian_operator(rng, dims, cplx)
          for dims, cplx in (([4, 5, 3], False), ([6, 7], True), ([3, 4, 2], True), ([17, 16], False))]
    m = 5
    singles = []
    for H in Hs:
        E, V = kfc.solve(kfc.Dense(), H, m)
        M = O.to_matrix(oracle_from_kfc_terms(H))
        assert np.allclose(M, M.conj().T, atol=1e-12)
        w = np.linalg.eigvalsh(M)[:m]
        assert np.max(np.abs(E - w)) <= 1e-10 * max(1.0, np.max(np.abs(w)))
        r = np.linalg.norm(M @ V - V * E, axis=0)
        assert np.max(r) <= 1e-9 * max(1.0, np.max(np.abs(w)))
        singles.append((E, V))
    batch, st = kfc.solve_dense_batch(Hs, m)
    for (Eb, Vb), (Es, Vs) in zip(batch, singles):
        assert np.array_equal(Eb, Es) and np.array_equal(Vb, Vs)
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_dense_batch(Hs, 0)
