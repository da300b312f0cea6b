"""Pins the CPU oracle: closed forms, brute-force Kronecker sums and the
reference's own golden energies (test/fixtures/scqubits_reference.json)."""
import json
import math
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

from oracle import kfc_oracle as O

RNG = np.random.default_rng(20261203)


def _ref(golden_dir, name):
    with open(os.path.join(golden_dir, "scqubits_reference.json")) as f:
        e = json.load(f)[name]
    return np.array(e["energies"]), e["params"]


# --- bases (test/test_bases.jl) ---------------------------------------------------------

def test_fd_weights_literal_and_moments():
    # the reference's literal second-derivative table (src/bases.jl:130-135)
    lit2 = {2: [1, -2, 1], 4: [Fr(-1, 12), Fr(4, 3), Fr(-5, 2), Fr(4, 3), Fr(-1, 12)],
            6: [Fr(1, 90), Fr(-3, 20), Fr(3, 2), Fr(-49, 18), Fr(3, 2), Fr(-3, 20), Fr(1, 90)],
            8: [Fr(-1, 560), Fr(8, 315), Fr(-1, 5), Fr(8, 5), Fr(-205, 72), Fr(8, 5), Fr(-1, 5), Fr(8, 315), Fr(-1, 560)]}
    for o in (2, 4, 6, 8):
        assert O._fd_derive(o, 2) == [Fr(x) for x in lit2[o]]
        assert O._fd_derive(o, 1) == [Fr(x) for x in O._FD1[o]]
        x = np.arange(-(o // 2), o // 2 + 1)
        w1, w2 = np.array(O.fd_weights(o, 1)), np.array(O.fd_weights(o, 2))
        assert abs(np.sum(w1 * x) - 1) < 1e-14 and abs(np.sum(w2)) < 1e-12
        assert abs(np.sum(w2 * x ** 2) - 2) < 1e-12
        assert np.allclose(w1, -w1[::-1]) and np.allclose(w2, w2[::-1])


def test_phase_grid_derivatives_on_gaussian():
    g = O.PhaseGrid(-10, 10, 801, 8)
    phi = g.coordinates()
    psi = np.exp(-phi ** 2 / 2)
    assert np.allclose(O.dense(O.localop(g, ("n2", 0))) @ psi, -(phi ** 2 - 1) * psi, atol=1e-7)
    N1 = O.dense(O.localop(g, ("n", 0)))
    assert np.allclose(N1 @ psi, 1j * phi * psi, atol=1e-7)
    assert np.allclose(N1, N1.conj().T)


def test_charge_and_fock_closed_forms():
    c = O.ChargeBasis(3)
    assert np.allclose(O.dense(O.localop(c, ("expiphi", 1))), np.diag(np.ones(6), -1))
    assert np.allclose(O.dense(O.localop(c, ("expiphi", -2))), np.diag(np.ones(5), 2))
    f = O.FockBasis(40, 0.7)
    a = np.diag(np.sqrt(np.arange(1, 40)), 1)
    nz = 1 / (2 * 0.7)
    assert np.allclose(O.dense(O.localop(f, ("phi", 0))), 0.7 * (a + a.T))
    assert np.allclose(O.dense(O.localop(f, ("n", 0))), 1j * nz * (a.T - a))
    n2 = (1j * nz * (a.T - a)) @ (1j * nz * (a.T - a))
    assert np.allclose(O.dense(O.localop(f, ("n2", 0)))[:38, :38], n2[:38, :38])
    assert np.isclose(O.dense(O.localop(f, ("n2", 0)))[39, 39], nz ** 2 * (2 * 39 + 1))


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_periodic_grid_plane_wave_symbols(order):
    """Free rotor on PeriodicGrid: 4E_C(sigma2(k) - 2 n_g sigma1(k) + n_g^2)
    (src/bases.jl:64-67; test/test_bases.jl:78-118)."""
    N, EC, ng = 24, 1.3, 0.31
    b = O.PeriodicGrid(N, order)
    D = b.spacing()
    h = order // 2
    w1, w2 = O.fd_weights(order, 1), O.fd_weights(order, 2)
    ks = np.arange(-N // 2, N // 2)
    s1 = np.array([sum(w1[j + h] * math.sin(j * k * D) for j in range(-h, h + 1)) / D for k in ks])
    s2 = np.array([-sum(w2[j + h] * math.cos(j * k * D) for j in range(-h, h + 1)) / D ** 2 for k in ks])
    Hm = 4 * EC * (O.dense(O.localop(b, ("n2", 0))) - 2 * ng * O.dense(O.localop(b, ("n", 0))) + ng ** 2 * np.eye(N))
    assert np.allclose(np.linalg.eigvalsh(Hm), np.sort(4 * EC * (s2 - 2 * ng * s1 + ng ** 2)), atol=1e-10)


# --- apply vs brute force (test/test_apply.jl, test/helpers.jl:10-24) --------------------

def _random_op(d, kind, cplx):
    def r(*s):
        x = RNG.standard_normal(s)
        return x + 1j * RNG.standard_normal(s) if cplx else x
    if kind == "dense":
        return O.DenseOp(r(d, d))
    if kind == "diag":
        return O.DiagonalOp(r(d))
    if kind == "banded":
        offs = [-2, 0, 1, d - 1]
        return O.BandedOp(d, offs, [r(d) for _ in offs])
    return O.ShiftOp(d, int(RNG.integers(-d + 1, d)), complex(*RNG.standard_normal(2)) if cplx else 0.7)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("dims", [[3, 4, 2], [5, 6, 4], [7]])
def test_apply_chain_vs_kron(dims, cplx):
    kinds = ["dense", "diag", "banded", "shift"]
    terms = []
    for t in range(6):
        axes = sorted(RNG.choice(len(dims), size=int(RNG.integers(1, len(dims) + 1)), replace=False))
        terms.append(O.AxisTerm(complex(*RNG.standard_normal(2)) if cplx else float(RNG.standard_normal()),
                                [(int(a), _random_op(dims[a], kinds[(t + a) % 4], cplx)) for a in axes]))
    H = O.kron_operator(dims, terms, constant=0.37)
    Hd = O.kron_reference(dims, terms, constant=0.37)
    X = RNG.standard_normal((H.N, 5))
    assert np.allclose(O.apply_chain(H, X), Hd @ X, atol=1e-12)
    assert np.allclose(O.to_matrix(H), Hd, atol=1e-12)
    assert np.allclose(O.full_diagonal(H), np.real(np.diag(Hd)))


def test_product_seed_and_gershgorin():
    b = O.PhaseGrid(-6 * np.pi, 6 * np.pi, 40, 2)
    H = O.three_fluxonia_operator([1.0, 0.9, 1.1], [0.5, 0.4, 0.6], [4, 5, 6], [0.1, 0.2, 0.3], N=12, order=2)
    Hd = O.to_matrix(H)
    assert np.max(np.linalg.eigvalsh(Hd)) <= O.gershgorin_bound(H) + 1e-12
    X = O.product_seed(H, 5)
    assert np.allclose(X.T @ X, np.eye(5))
    assert b.dim == 40


# --- end to end vs the reference's golden energies ----------------------------------------

def test_transmon_vs_scqubits(golden_dir):
    E, p = _ref(golden_dir, "transmon")
    root = O.Node([("varphi_1", O.ChargeBasis(30))], keep=5)
    mat = O.materialize(root, O.transmon_spec(p["E_J"], p["E_C"], p["n_g"]))
    assert np.max(np.abs(mat.energies - E)) < 1e-10


def test_fluxonium_vs_scqubits(golden_dir):
    E, p = _ref(golden_dir, "fluxonium")
    for real in (True, False):
        root = O.Node([("varphi_1", O.PhaseGrid(-10 * np.pi, 10 * np.pi, 1201, 8))], keep=6, real=real)
        mat = O.materialize(root, O.fluxonium_spec(p["E_J"], p["E_C"], p["E_L"], p["varphi_ext"]))
        assert np.max(np.abs(mat.energies - E)) < 1e-6


def test_zero_pi_small_vs_scqubits(golden_dir):
    E, p = _ref(golden_dir, "zero_pi_small")
    spec = O.zero_pi_spec(p["E_J"], p["E_L"], p["E_CJ"], p["E_CS"], p["varphi_ext"], p["n_g"])
    for real in (True, False):
        root = O.Node([("theta", O.ChargeBasis(6)), ("phi", O.PhaseGrid(-6 * np.pi, 6 * np.pi, 101, 6))], keep=8, real=real)
        mat = O.materialize(root, spec)
        assert np.max(np.abs(mat.energies - E)) < 1e-9


def test_full_zero_pi_small_vs_scqubits(golden_dir):
    E, p = _ref(golden_dir, "full_zero_pi_small")
    spec = O.full_zero_pi_spec(p["E_J"], p["E_L"], p["E_CJ"], p["E_CS"], p["E_Cz"], p["dE_J"], p["dC_J"], p["dC"],
                               p["dE_L"], p["varphi_ext"], p["n_g"])
    zp = O.Node([("theta", O.ChargeBasis(6)), ("phi", O.PhaseGrid(-6 * np.pi, 6 * np.pi, 101, 6))], keep=30, real=False)
    zeta = O.Node([("zeta", O.FockBasis(20, (2 * p["E_Cz"] / p["E_L"]) ** 0.25))], keep=20)
    full = O.Node([zp, zeta], keep=10, real=False)
    mat = O.materialize(full, spec)
    assert np.max(np.abs(mat.energies - E)) < 1e-9


def test_two_transmons_vs_scqubits(golden_dir):
    """test/test_materialize.jl:142-160: tree lowest 3 to 1e-8, all 20 to 5e-4,
    scqubits' bare construction to 1e-10."""
    E, p = _ref(golden_dir, "two_transmons")
    spec = O.two_transmons_spec(p["E_J1"], p["E_C1"], p["n_g1"], p["E_J2"], p["E_C2"], p["n_g2"], p["g"])
    top = O.Node([O.Node([("varphi_1", O.ChargeBasis(30))], keep=8), O.Node([("varphi_2", O.ChargeBasis(30))], keep=8)],
                 keep=20)
    mat = O.materialize(top, spec)
    assert np.max(np.abs(mat.energies[:3] - E[:3])) < 1e-8
    assert np.max(np.abs(mat.energies - E)) < 5e-4

    def bare(EJ, EC, ng):
        m = O.materialize(O.Node([("varphi_1", O.ChargeBasis(30))], keep=8), O.transmon_spec(EJ, EC, ng))
        return m, O.port_matrix(m, "varphi_1", ("n", 0.0))
    b1, n1 = bare(p["E_J1"], p["E_C1"], p["n_g1"])
    b2, n2 = bare(p["E_J2"], p["E_C2"], p["n_g2"])
    I8 = np.eye(8)
    Hsc = np.kron(I8, np.diag(b1.energies)) + np.kron(np.diag(b2.energies), I8) + \
        p["g"] * np.kron(n2 - p["n_g2"] * I8, n1 - p["n_g1"] * I8)
    assert np.max(np.abs(np.linalg.eigvalsh(Hsc)[:20] - E)) < 1e-10


def test_real_expansion_equals_complex_three_fluxonia():
    args = ([1.0, 0.9, 1.1], [0.5, 0.4, 0.6], [4.0, 5.0, 6.0], [0.1, 2.2, 4.3])
    Hr = O.three_fluxonia_operator(*args, N=9, order=2, real=True)
    Hc = O.three_fluxonia_operator(*args, N=9, order=2, real=False)
    Mr, Mc = O.to_matrix(Hr), O.to_matrix(Hc)
    assert Hr.dtype == np.float64 and Hc.dtype == np.complex128
    assert np.allclose(Mr, Mc, atol=1e-12)
    assert len(Hr.terms) == 6      # 3 banded n^2 + 3 coupling terms after fusion of diagonal ones


def test_mlp_backward_by_finite_differences():
    rng = np.random.default_rng(5)
    shapes = [(7, 6), (7,), (5, 7), (5,), (4, 5), (4,)]
    params = [rng.standard_normal(s) * 0.5 for s in shapes]
    X = rng.standard_normal((3, 6))
    dy = rng.standard_normal((3, 4))
    y, cache = O.mlp_forward(params, X)
    grads = O.mlp_backward(params, cache, dy)
    for pi, g in zip(range(6), grads):
        for idx in [tuple(0 for _ in shapes[pi]), tuple(s - 1 for s in shapes[pi])]:
            h = 1e-6
            pp = [q.copy() for q in params]
            pp[pi][idx] += h
            yp, _ = O.mlp_forward(pp, X)
            pp[pi][idx] -= 2 * h
            ym, _ = O.mlp_forward(pp, X)
            fd = np.sum(dy * (yp - ym)) / (2 * h)
            assert abs(fd - g[idx]) < 1e-6 * max(1, abs(fd))


def test_c_restatement_equals_numpy_chain():
    from oracle import c_apply
    if not c_apply.available():
        pytest.skip("oracle/_build/liboracle_apply.so not built")
    rng = np.random.default_rng(9)
    H = O.three_fluxonia_operator([1.0, 0.9, 1.1], [0.5, 0.4, 0.6], [4, 5, 6], [0.1, 0.2, 0.3], N=11, order=2)
    x = rng.standard_normal(H.N)
    assert np.allclose(c_apply.CApply(H)(x), O.apply_chain(H, x), atol=1e-12)
    dims = [4, 5, 3]
    terms = [O.AxisTerm(0.7, [(0, O.DenseOp(rng.standard_normal((4, 4)))), (2, O.ShiftOp(3, 1, 0.5))]),
             O.AxisTerm(-1.1, [(1, O.BandedOp(5, [-2, 1], [rng.standard_normal(5), rng.standard_normal(5)]))])]
    H2 = O.kron_operator(dims, terms, 0.2)
    x2 = rng.standard_normal(H2.N)
    assert np.allclose(c_apply.CApply(H2)(x2), O.apply_chain(H2, x2), atol=1e-12)


# --- the linear solve on the complement (src/resolvent.jl:119-256) -----------------------

def _spectral_case(n=60, k=4, cplx=False, seed=3):
    r = np.random.default_rng(seed)
    A = r.standard_normal((n, n)) + (1j * r.standard_normal((n, n)) if cplx else 0)
    H = (A + A.conj().T) / 2 + np.diag(np.linspace(0, 40, n))
    E, U = np.linalg.eigh(H)
    return H, E, U


@pytest.mark.parametrize("cplx", [False, True])
def test_complement_solve_spectral_solution(cplx):
    """Sternheimer columns: basis = the k lowest states, shift s_j = E_j, right-hand
    side a random vector; the exact solution on the complement is
    sum_{l >= k} psi_l <psi_l|b_j> / (E_l - s_j)."""
    H, E, U = _spectral_case(cplx=cplx)
    k, n = 4, 4
    r = np.random.default_rng(7)
    Bm = r.standard_normal((H.shape[0], n)) + (1j * r.standard_normal((H.shape[0], n)) if cplx else 0)
    shifts = E[:n]
    X, it, res, fs, af, fr = O.complement_solve(lambda Y: H @ Y, np.real(np.diag(H)), U[:, :k], Bm, shifts,
                                                tol=1e-12, maxiter=500)
    Uo = U[:, k:]
    Xe = Uo @ ((Uo.conj().T @ Bm) / (E[k:, None] - shifts[None, :]))
    assert np.max(np.abs(X - Xe)) <= 1e-9 * np.max(np.abs(Xe))
    assert np.max(np.abs(U[:, :k].conj().T @ X)) <= 1e-12 * np.max(np.abs(X))
    assert 0 < it < 200 and fs == [] and np.all(res <= 10 * 1e-12)


def test_complement_solve_from_start_and_errors():
    H, E, U = _spectral_case()
    k = 4
    Bm = np.zeros((H.shape[0], 3))
    Bm[:, 1] = U[:, 10]               # on the complement
    Bm[:, 2] = U[:, 0]                # inside the basis: nothing on the complement
    X, it, res, fs, af, fr = O.complement_solve(lambda Y: H @ Y, np.diag(H), U[:, :k], Bm, E[:3],
                                                tol=1e-10, maxiter=500, scale=np.array([1.0, 1.0, 1.0]))
    assert fs == [0, 2] and np.all(X[:, [0, 2]] == 0)
    assert np.allclose(X[:, 1], U[:, 10] / (E[10] - E[1]), atol=1e-9)
    # a shift above a state outside the basis: H - s is not positive on the complement
    with pytest.raises(O.ComplementError, match="not positive on the complement"):
        O.complement_solve(lambda Y: H @ Y, np.diag(H), U[:, :2], U[:, [6]] + U[:, [3]], [E[5]], tol=1e-10,
                           maxiter=500)
    with pytest.raises(O.ComplementError, match="did not converge"):
        O.complement_solve(lambda Y: H @ Y, np.diag(H), U[:, :k], np.ones((H.shape[0], 1)), [E[0]], tol=1e-14,
                           maxiter=3)
