"""Lockstep LOBPCG over a sweep (kfc_lobpcg_lockstep; no reference counterpart for
the batching, each point keeps the contract of lobpcg_solve, src/lobpcg.jl:217-301):
every point's pairs match the one-point path, with orthonormal vectors and fresh
FP64 residuals within the tolerance."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_sweep():
    from kfc.sweeps import ZeroPiSweep
    return ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, contexts=2)


def _check_point(H, E, V, Eref, tol, rel):
    import kfc
    assert np.max(np.abs(E - Eref) / np.abs(Eref)) < rel
    assert np.allclose(V.T @ V, np.eye(V.shape[1]), atol=1e-9)
    assert np.max(kfc.residuals(H, E, V)) <= tol


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_lockstep_matches_sequential(mode):
    from kfc import configs
    sw = _small_sweep()
    pts = [configs.zero_pi_theta(7100 + i) for i in range(7)]
    seq, _ = sw.run_plain(pts)
    lock, _ = sw.run_batched(pts, mode=mode, lockstep=True)
    for (Es, Vs, ss, _), (El, Vl, sl, _) in zip(seq, lock):
        assert sl["max_residual"] <= 1e-8
        assert sl["iterations"] > 0 and sl["n_matvec"] > 0
        if mode == "mixed":
            assert 0 < sl["n_matvec_f32"] < sl["n_matvec"]
    Hs = [sw.parent(th)[0] for th in pts]
    for H, (Es, _, _, _), (El, Vl, _, _) in zip(Hs, seq, lock):
        _check_point(H, El, Vl, Es, 1e-8, 1e-9)


def test_lockstep_full_parent():
    """Four points of the zero-pi parent [40, 40, 30] (48 000 states, m = 16)."""
    from kfc import configs
    from kfc.sweeps import ZeroPiSweep
    sw = ZeroPiSweep(contexts=1)
    pts = [configs.zero_pi_theta(8100 + i) for i in range(4)]
    seq, _ = sw.run_plain(pts[:2])
    lock, _ = sw.run_batched(pts, mode="f64", lockstep=True)
    for (Es, _, _, _), (El, Vl, sl, _) in zip(seq, lock):
        assert sl["max_residual"] <= 1e-8
        assert np.max(np.abs(Es - El) / np.abs(Es)) < 1e-9
    for th, (El, Vl, sl, _) in zip(pts, lock):
        H, _ = sw.parent(th)
        assert np.allclose(Vl.T @ Vl, np.eye(Vl.shape[1]), atol=1e-9)
        assert np.max(__import__("kfc").residuals(H, El, Vl)) <= 1e-8


def test_lockstep_warm_start():
    """X0 = converged vectors of the same points, guard columns included (the
    warm-guards seed): a few iterations instead of about a hundred, and the same
    iteration counts as the one-point path from the same blocks."""
    from kfc import configs
    sw = _small_sweep()
    pts = [configs.zero_pi_theta(7300 + i) for i in range(5)]
    seq, _ = sw.run_plain(pts)
    W = sw.width(12 * 12 * 10)
    full, _ = sw.run_batched(pts, mode="f64", lockstep=True, nev=W)
    cold, _ = sw.run_batched(pts, mode="f64", lockstep=True)
    warm, _ = sw.run_batched(pts, mode="f64", lockstep=True, x0_fn=lambda th, Hs, ks: [o[1] for o in full])
    for (Es, _, _, _), (Ew, Vw, sw_, _), (_, _, sc, _) in zip(seq, warm, cold):
        assert sw_["warm"] == 1 and sw_["max_residual"] <= 1e-8
        assert np.max(np.abs(Es - Ew) / np.abs(Es)) < 1e-9
        assert sw_["iterations"] <= 10 < sc["iterations"]


def test_lockstep_stencil_operators():
    """Operators off the dense-pair path (fused stencils): one apply launch per
    point inside each lockstep step."""
    import kfc
    from kfc import configs
    Hs = [configs.three_fluxonia_operator(configs.three_fluxonia_theta(s), N=12) for s in (1, 2, 3)]
    s = kfc.LOBPCG(tol=1e-8, mode="f64")
    res, st = kfc.solve_batch(s, Hs, 4, lockstep=True)
    for H, (E, V), x in zip(Hs, res, st):
        Ed, _ = kfc.solve(kfc.Dense(), H, 4)
        assert x["max_residual"] <= 1e-8
        _check_point(H, E, V, Ed, 1e-8, 1e-9)


def test_lockstep_rejects():
    import kfc
    from kfc import configs
    from kfc import _lib as L
    s = kfc.LOBPCG(tol=1e-8)
    H1 = configs.three_fluxonia_operator(configs.three_fluxonia_theta(1), N=12)
    H2 = configs.three_fluxonia_operator(configs.three_fluxonia_theta(2), N=10)
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch(s, [H1, H2], 4, lockstep=True)          # sizes differ
    H3 = configs.three_fluxonia_operator(configs.three_fluxonia_theta(3), N=12, ctx=L.Context(0))
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch(s, [H1, H3], 4, lockstep=True)          # two contexts
    with pytest.raises(kfc.UnsupportedError):
        kfc.solve_batch(kfc.LOBPCG(tol=1e-8, guard=4), [H1], 18, lockstep=True)   # 3 (m + g) > 64
