"""Matrix-free time evolution on a node's product basis: port_operator
(src/portoperator.jl:22-39) and its consumers KronFreeApply / InteractionApply /
drive_hamiltonian (ext/KronFreeCircuitsQuantumToolboxExt.jl:44-87, 185-280).

Pinned by: the port operator's apply against the oracle's dense matrix of the same
terms, and its projection V'OV onto the node's states against the node's port (the
host formula, itself pinned against scqubits through the tree energies); the two
applies against dense numpy formulas; and the evolutions against exp(-iHt) (no
drive) and against each other (lab frame = interaction frame rotated back)."""
import math

import numpy as np
import pytest

from helpers import oracle_from_kfc_terms, random_ops, to_kfc, to_oracle
from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zero_pi_small():
    import kfc
    from kfc import configs
    th = configs.zero_pi_theta(4400)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    tree = configs.zero_pi_tree(E_Cz, ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, solver=kfc.Dense())
    return kfc.materialize(tree, spec).root


def _dense(H):
    return O.to_matrix(oracle_from_kfc_terms(H))


def _ket(rng, N, b=None, device=True):
    import torch
    shape = (N,) if b is None else (b, N)
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    return torch.from_numpy(z).to("cuda") if device else z


KEYS = [("theta", ("n", 0)), ("theta", ("expiphi", 1)), ("phi", ("phi", 1)), ("zeta", ("n", 0)),
        [("theta", ("expiphi", 1)), ("phi", ("expiphi", 1))], [("theta", ("n", 0)), ("zeta", ("n", 0))]]


@pytest.mark.parametrize("key", KEYS, ids=[str(k) for k in KEYS])
def test_port_operator_parent(zero_pi_small, key):
    import torch
    from kfc.tree import joint_port, port_matrix, port_operator
    mn = zero_pi_small
    Op = port_operator(mn, key)
    M = _dense(Op)
    rng = np.random.default_rng(3)
    X = _ket(rng, Op.N, 3, device=False)
    if Op.dtype != np.complex128:
        X = np.ascontiguousarray(X.real)
    Y = Op @ torch.from_numpy(X).to("cuda")
    Op.sync()
    assert np.allclose(Y.cpu().numpy(), X @ M.T, rtol=0, atol=1e-12 * max(1.0, np.abs(M).sum(1).max()))
    V = mn.vectors
    P = port_matrix(mn, *key, device=False) if isinstance(key, tuple) else joint_port(mn, key, device=False)
    assert np.allclose(V.conj().T @ (M @ V), P, rtol=0, atol=1e-10 * max(1.0, np.abs(P).max()))
    assert Op.hermitian == bool(np.allclose(M, M.conj().T, atol=1e-12 * max(1.0, np.abs(M).max())))


def test_port_operator_basis_children():
    """A flat node: the factors are the leaf bases' own operators."""
    import kfc
    from kfc import configs
    from kfc.tree import port_operator
    th = configs.three_fluxonia_theta(2)
    spec = configs.three_fluxonia_spec(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"])
    mn = kfc.materialize(configs.three_fluxonia_node(N=10, m=4, solver=kfc.Dense()), spec).root
    for key in (("phi2", ("n", 0)), [("phi1", ("phi", 1)), ("phi3", ("n", 0))], ("phi1", ("expiphi", 1))):
        Op = port_operator(mn, key)
        bases = {m: mn.node.leaf_basis(m) for m in ("phi1", "phi2", "phi3")}
        keys = [key] if isinstance(key, tuple) else key
        mats = [np.eye(10)] * 3
        for m, p in keys:   # the leaf operators themselves are pinned in tests/test_oracle.py
            mats[["phi1", "phi2", "phi3"].index(m)] = kfc.dense(kfc.localop(bases[m], p))
        ref = np.kron(mats[2], np.kron(mats[1], mats[0]))
        assert np.allclose(_dense(Op), ref, atol=1e-12)


def test_port_operator_rejects(zero_pi_small):
    import kfc
    from kfc.tree import port_operator
    with pytest.raises(kfc.ArgumentError):
        port_operator(zero_pi_small, [])
    with pytest.raises(kfc.ArgumentError):
        port_operator(zero_pi_small, ("chi", ("n", 0)))
    with pytest.raises(kfc.ArgumentError):
        port_operator(zero_pi_small, [("theta", ("n", 0)), ("theta", ("expiphi", 1))])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("b", [None, 3])
def test_kronfree_apply_vs_dense(cplx, b):
    import kfc
    rng = np.random.default_rng(31 + cplx)
    dims = [5, 4, 6]
    terms = random_ops(rng, dims, 5, cplx)
    H = to_kfc(dims, terms, constant=0.4)
    M = O.to_matrix(to_oracle(dims, terms, 0.4))
    A = kfc.KronFreeApply(H, 2 * math.pi, 0.7)
    v = _ket(rng, H.N, b)
    w = _ket(rng, H.N, b)
    w0 = w.cpu().numpy().copy()
    vv = v.cpu().numpy()
    R = lambda x: (x @ M.T) if x.ndim == 2 else (M @ x)
    for alpha, beta in ((1.0, 0.0), (0.3 - 0.8j, 0.0), (-1j, 0.5 + 0.25j)):
        ref = alpha * 2 * math.pi * (R(vv) - 0.7 * vv) + (beta * w0 if beta != 0 else 0)
        wt = w.clone()
        A(wt, v, alpha, beta)
        got = wt.cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("b", [None, 4])
def test_kronfree_apply_stencil_operator(b):
    """A real operator on the fused stencil path (three fluxonia, 10^3 states): the
    stencil runs on complex kets directly."""
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(4)
    H = configs.three_fluxonia_operator(th, N=10)
    assert H.fused
    M = O.to_matrix(O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=10, order=2))
    rng = np.random.default_rng(71)
    v, w = _ket(rng, H.N, b), _ket(rng, H.N, b)
    vv, w0 = v.cpu().numpy(), w.cpu().numpy()
    R = lambda x: (x @ M.T) if x.ndim == 2 else (M @ x)
    A = kfc.KronFreeApply(H, 2 * math.pi, -2.0)
    for alpha, beta in ((1.0, 0.0), (-1j, 0.0), (0.4 + 0.3j, 1.0), (-1j, 0.2 - 0.9j)):
        ref = alpha * 2 * math.pi * (R(vv) + 2.0 * vv) + (beta * w0 if beta != 0 else 0)
        wt = w.clone()
        A(wt, v, alpha, beta)
        got = wt.cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    with pytest.raises(kfc.ArgumentError):
        A(v, v)
    # the interaction frame of the same operator (the diagonal subtraction rides in
    # the combination after the stencil)
    IA = kfc.InteractionApply(H, -2.0, [])
    Om = 2 * math.pi * (np.diag(M) + 2.0)
    t = 0.37
    ph = np.exp(-1j * Om * t)
    Hb = 2 * math.pi * (M + 2.0 * np.eye(H.N)) - np.diag(Om)
    wt = w.clone()
    IA(wt, v, t, None, 1.0, 0.5)
    if vv.ndim == 1:
        ref = np.conj(ph) * (Hb @ (ph * vv)) + 0.5 * w0
    else:
        ref = np.conj(ph)[None, :] * ((ph[None, :] * vv) @ Hb.T) + 0.5 * w0
    assert np.max(np.abs(wt.cpu().numpy() - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_interaction_apply_vs_dense():
    import kfc
    import torch
    rng = np.random.default_rng(41)
    dims = [4, 5, 3]
    H0 = to_kfc(dims, random_ops(rng, dims, 4, False), constant=1.5)
    t_real = random_ops(rng, dims, 2, False)
    t_cplx = random_ops(rng, dims, 2, True)
    O1, O2 = to_kfc(dims, t_real), to_kfc(dims, t_cplx)
    M0 = O.to_matrix(oracle_from_kfc_terms(H0))
    M1, M2 = O.to_matrix(to_oracle(dims, t_real)), O.to_matrix(to_oracle(dims, t_cplx))
    E0, t = -0.35, 0.123
    f1 = lambda p, t: 0.2 * math.cos(3 * t)
    f2 = lambda p, t: 0.1j * math.sin(2 * t) + 0.05
    A = kfc.InteractionApply(H0, E0, [(O1, f1), (O2, f2)])
    Om = 2 * math.pi * (np.real(np.diag(M0)) - E0)
    assert np.allclose(A.omega.cpu().numpy(), Om, atol=1e-12)
    v = _ket(rng, H0.N, 2)
    w = _ket(rng, H0.N, 2)
    vv, w0 = v.cpu().numpy(), w.cpu().numpy()
    ph = np.exp(-1j * Om * t)
    Hb = 2 * math.pi * (M0 - E0 * np.eye(H0.N)) - np.diag(Om) + 2 * math.pi * (f1(None, t) * M1 + f2(None, t) * M2)
    alpha, beta = 0.7 - 0.2j, -0.3j
    ref = alpha * np.conj(ph)[None, :] * ((ph[None, :] * vv) @ Hb.T) + beta * w0
    A(w, v, t, None, alpha, beta)
    got = w.cpu().numpy()
    assert np.max(np.abs(got - ref)) <= 1e-11 * np.max(np.abs(ref))
    with pytest.raises(kfc.DimensionMismatch):
        kfc.InteractionApply(H0, E0, [(to_kfc([4, 5, 2], random_ops(rng, [4, 5, 2], 1, False)), f1)])


def _steps(Hd, t1, p=None):
    # RK4 with h |2 pi (H0 - E0)| <= 0.02, the norm bounded by Gershgorin
    r = 2 * math.pi * abs(Hd.H0.gershgorin_bound() - Hd.E0)
    return int(math.ceil(t1 * r / 0.02))


def test_evolution_no_drive_matches_expm(zero_pi_small):
    import scipy.linalg
    import torch
    import kfc
    mn = zero_pi_small
    Hd = kfc.DriveHamiltonian(mn, [])
    M0 = _dense(mn.operator)
    rng = np.random.default_rng(51)
    psi0 = _ket(rng, mn.operator.N, device=False)
    psi0 /= np.linalg.norm(psi0)
    t1 = 0.01
    psi = kfc.evolve_rk4(Hd, torch.from_numpy(psi0).to("cuda"), 0.0, t1, _steps(Hd, t1))
    ref = scipy.linalg.expm(-1j * 2 * math.pi * t1 * (M0 - Hd.E0 * np.eye(M0.shape[0]))) @ psi0
    assert np.linalg.norm(psi.cpu().numpy() - ref) <= 1e-7
    assert abs(np.linalg.norm(psi.cpu().numpy()) - 1.0) <= 1e-7


def test_lab_and_interaction_frames_agree(zero_pi_small):
    """A driven evolution in the lab frame equals the interaction-frame evolution
    rotated back by e^{-i Omega t} (both started from the same state at t = 0), and
    keeps the norm (Hermitian ports, real drive coefficients)."""
    import torch
    import kfc
    mn = zero_pi_small
    drives = [(("phi", ("phi", 1)), lambda p, t: 0.8 * math.cos(2 * math.pi * 1.3 * t)),
              ([("theta", ("n", 0)), ("zeta", ("n", 0))], lambda p, t: 0.3 * math.sin(2 * math.pi * 0.7 * t))]
    lab = kfc.DriveHamiltonian(mn, drives, frame="lab")
    itr = kfc.DriveHamiltonian(mn, drives, frame="interaction")
    rng = np.random.default_rng(61)
    psi0 = _ket(rng, mn.operator.N, device=False)
    psi0 = torch.from_numpy(psi0 / np.linalg.norm(psi0)).to("cuda")
    t1 = 0.01
    n = _steps(lab, t1) + 50
    psi_l = kfc.evolve_rk4(lab, psi0, 0.0, t1, n)
    psi_i = kfc.evolve_rk4(itr, psi0, 0.0, t1, n)
    back = itr.to_lab(psi_i, t1)
    assert torch.linalg.norm(psi_l - back).item() <= 1e-6
    assert abs(torch.linalg.norm(psi_l).item() - 1.0) <= 1e-6
