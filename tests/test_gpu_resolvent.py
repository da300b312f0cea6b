"""The linear solve on the complement of a node's known states on the device
(kfc_complement_solve, _complement_solve of src/resolvent.jl:146-256) and the
resolvent built on it (_apply_resolvent, :270-331).

Pinned by: the oracle restatement (tests/test_oracle.py pins it against the spectral
solution on dense matrices) on the same inputs -- same iterations and flags, solutions
within the tolerance -- and, end to end, the Sternheimer derivative of an eigenvector
against central finite differences of the Dense eigenvectors."""
import numpy as np
import pytest

from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu


def _fluxonia(N=8, seed=3, **shift):
    from kfc import configs
    th = configs.three_fluxonia_theta(seed)
    for key, (i, d) in shift.items():
        th[key] = list(th[key])
        th[key][i] += d
    return th, configs.three_fluxonia_operator(th, N=N)


def _dense_of(th, N):
    return O.to_matrix(O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=N, order=2))


def _transmon_top():
    from kfc import transmon_sweep as T
    from kfc.tree import operator
    inst = T.draw(1, seed=11)[0]
    inst["N1"], inst["N2"] = 128, 128
    H, _ = operator(T.tree_of(inst, k_leaf=8, k_top=20), T.spec_of(inst))
    return H


def _dense_kfc(H):
    import torch
    I = torch.eye(H.N, dtype=torch.complex128 if H.dtype == np.complex128 else torch.float64, device="cuda")
    Y = H @ I
    H.sync()
    return Y.cpu().numpy().T


@pytest.mark.parametrize("case", ["real", "complex"])
def test_complement_solve_matches_oracle(case):
    import kfc
    if case == "real":
        th, H = _fluxonia()
        M = _dense_of(th, 8)
    else:
        H = _transmon_top()
        M = _dense_kfc(H)
    E, U = np.linalg.eigh(M)
    k, n = 4, 4
    rng = np.random.default_rng(5)
    B = rng.standard_normal((H.N, n)) + (1j * rng.standard_normal((H.N, n)) if case == "complex" else 0)
    B[:, 3] = U[:, 1]             # inside the basis: converged from the start
    D = np.real(np.diag(M))
    tol = 1e-10
    Xo, ito, reso, fso, afo, fro = O.complement_solve(lambda Y: M @ Y, D, U[:, :k], B, E[:n], tol=tol, maxiter=800)
    X, it, res, fs, af, fr = kfc.complement_solve(H, H.diagonal_device(), U[:, :k].T, B.T, E[:n], tol, 800)
    X = X.cpu().numpy().T
    assert fs == fso == [3] and af == afo
    assert abs(it - ito) <= 1
    assert np.all(X[:, 3] == 0)
    assert np.max(np.abs(X - Xo)) <= 1e-8 * np.max(np.abs(Xo))
    Xe = U[:, k:] @ ((U[:, k:].conj().T @ B) / (E[k:, None] - E[None, :n]))
    assert np.max(np.abs(X - Xe)) <= 1e-7 * np.max(np.abs(Xe))
    assert np.all(res[:3] <= 10 * tol)


def test_complement_solve_errors():
    import kfc
    th, H = _fluxonia()
    M = _dense_of(th, 8)
    E, U = np.linalg.eigh(M)
    D = H.diagonal_device()
    with pytest.raises(kfc.BreakdownError, match="not positive on the complement"):
        kfc.complement_solve(H, D, U[:, 2:4].T, (U[:, 6] + U[:, 0])[None, :], [E[4]], 1e-10, 500)
    with pytest.raises(kfc.ConvergenceError, match="did not converge"):
        kfc.complement_solve(H, D, U[:, :4].T, np.ones((1, H.N)), [E[0]], 1e-13, 2)
    with pytest.raises(kfc.ArgumentError):
        kfc.complement_solve(H, D, U[:, :4].T, np.ones((1, H.N)), [E[0]], -1.0, 10)


def test_many_columns_in_groups():
    """70 right-hand sides: solved 64 and 6 at a time, each column to the same answer."""
    import kfc
    th, H = _fluxonia()
    M = _dense_of(th, 8)
    E, U = np.linalg.eigh(M)
    rng = np.random.default_rng(9)
    B = rng.standard_normal((H.N, 70))
    s = np.full(70, E[0])
    X, it, res, fs, af, fr = kfc.complement_solve(H, H.diagonal_device(), U[:, :3].T, B.T, s, 1e-10, 800)
    Xe = U[:, 3:] @ ((U[:, 3:].T @ B) / (E[3:, None] - E[0]))
    assert np.max(np.abs(X.cpu().numpy().T - Xe)) <= 1e-7 * np.max(np.abs(Xe))


def test_sternheimer_derivative_vs_finite_differences():
    """d psi_n / d E_J1 = sum_{l != n} psi_l <psi_l|dH psi_n> / (E_n - E_l): the kept
    states' part by the sum, the rest by the complement solve, against central
    differences of the Dense eigenvectors (10^3 states, 4 kept)."""
    import kfc
    import torch
    from kfc.resolvent import Complement, Resolvent, apply_resolvent
    N, k, h = 10, 4, 1e-4
    th, H = _fluxonia(N)
    _, Hp = _fluxonia(N, E_J=(0, h))
    _, Hm = _fluxonia(N, E_J=(0, -h))
    E, V = kfc.solve(kfc.Dense(), H, k)
    _, Vp = kfc.solve(kfc.Dense(), Hp, k)
    _, Vm = kfc.solve(kfc.Dense(), Hm, k)
    for j in range(k):   # align the signs of the perturbed vectors
        Vp[:, j] *= np.sign(Vp[:, j] @ V[:, j])
        Vm[:, j] *= np.sign(Vm[:, j] @ V[:, j])
    Vt = torch.from_numpy(np.ascontiguousarray(V.T)).to("cuda")
    dH = ((Hp @ Vt) - (Hm @ Vt)) / (2 * h)
    Hp.sync(), Hm.sync()
    R = dH.cpu().numpy().T
    res = Resolvent(np.asarray(E), V, Complement(H, H.diagonal_device(), Vt, 1e-11, 2000), 1e-9)
    W, clashes = apply_resolvent(res, R, E, [[j] for j in range(k)])
    assert all(c == [] for c in clashes)
    fd = (Vp - Vm) / (2 * h)
    # derivative of a normalized real eigenvector: orthogonal to it, the sum over l != n
    assert np.max(np.abs(W - fd)) <= 2e-5 * np.max(np.abs(fd))


def test_round_off_floor_columns():
    """A right-hand side almost inside the basis: its complement part (1e-11 of it)
    cannot reach tol on its own scale, so the column stops at the round-off floor
    (src/resolvent.jl:128-140), as the oracle's does; a column met by its tolerance
    beside it is unaffected."""
    import kfc
    th, H = _fluxonia()
    M = _dense_of(th, 8)
    E, U = np.linalg.eigh(M)
    k = 4
    B = np.stack([U[:, 0] + 1e-11 * U[:, 7], U[:, 9]], axis=1)
    args = dict(tol=1e-10, maxiter=800)
    Xo, ito, reso, fso, afo, fro = O.complement_solve(lambda Y: M @ Y, np.diag(M), U[:, :k], B, E[:2], **args)
    X, it, res, fs, af, fr = kfc.complement_solve(H, H.diagonal_device(), U[:, :k].T, B.T, E[:2], args["tol"],
                                                  args["maxiter"])
    assert afo == [0] and af == [0] and fs == fso == []
    assert 0 < fr[0] <= 1000 * np.finfo(float).eps and res[0] == 0
    assert res[1] <= 1e-9 and fr[1] == 0
    X = X.cpu().numpy().T
    assert np.allclose(X[:, 1], U[:, 9] / (E[9] - E[1]), atol=1e-9)
    assert np.max(np.abs(X[:, 0] - Xo[:, 0])) <= 1e-12
