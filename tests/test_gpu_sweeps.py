"""The learned warm start on a small zero-pi sweep (no reference counterpart: its
output is LOBPCG's X0, src/lobpcg.jl:217).  Trained on eigenpairs from the plain
solver, its predicted blocks start LOBPCG closer to the answer than the product
seed: fewer iterations, the same energies and the same residual bar."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_sweep():
    from kfc.sweeps import ZeroPiSweep
    return ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, contexts=2)


def test_warm_start_reduces_iterations():
    from kfc import configs
    from kfc.sweeps import WarmStart, subspace_capture
    sw = _small_sweep()
    train = [configs.zero_pi_theta(i) for i in range(48)]
    test = [configs.zero_pi_theta(9000 + i) for i in range(6)]
    nev = sw.width(12 * 12 * 10)
    gen, _ = sw.run_batched(train, mode="f64", nev=nev)
    assert all(o[2]["max_residual"] <= 1e-8 for o in gen)
    frames = [sw.to_frame(V, sw.frame_index(kids)) for E, V, st, kids in gen]
    ws = WarmStart(sw, nbasis=48, seed=3)
    loss0, loss1 = ws.train(train, frames, steps=1500, lr=1e-3)
    assert loss1 < 0.2 * loss0
    cold, _ = sw.run_batched(test, mode="f64")
    warm, _ = sw.run_batched(test, mode="f64", x0_fn=ws.x0_fn)
    it_cold = np.mean([o[2]["iterations"] for o in cold])
    it_warm = np.mean([o[2]["iterations"] for o in warm])
    for (Ec, Vc, sc, _), (Ew, Vw, sw_, _) in zip(cold, warm):
        assert sc["max_residual"] <= 1e-8 and sw_["max_residual"] <= 1e-8
        assert np.max(np.abs(Ec - Ew) / np.abs(Ec)) < 1e-9
        assert sw_["warm"] == 1
    # the predicted block holds most of each true eigenvector
    caps = []
    for xf, (E, V, st, kids) in zip(ws.predict_frames(test), cold):
        caps.append(np.mean(subspace_capture(sw.from_frame(xf, sw.frame_index(kids), V.shape[0]), V[:, :sw.m])))
    assert np.mean(caps) > 0.9
    assert it_warm < 0.8 * it_cold, (it_warm, it_cold)


def test_batched_matches_sequential():
    from kfc import configs
    sw = _small_sweep()
    pts = [configs.zero_pi_theta(7000 + i) for i in range(5)]
    seq, _ = sw.run_plain(pts)
    bat, _ = sw.run_batched(pts, mode="mixed")
    for (Es, Vs, ss, _), (Eb, Vb, sb, _) in zip(seq, bat):
        assert sb["max_residual"] <= 1e-8
        assert np.max(np.abs(Es - Eb) / np.abs(Es)) < 1e-8
        assert np.allclose(Vb.T @ Vb, np.eye(Vb.shape[1]), atol=1e-10)


def test_transmon_batch_vs_oracle():
    """Coupled transmons on periodic grids (complex nodes): the batched leaves
    (mixed precision, several contexts) and the plain path give the oracle's
    dense tree energies."""
    from oracle import kfc_oracle as O
    from kfc import transmon_sweep as T
    insts = T.draw(4, seed=11)
    for i, N in enumerate((64, 48, 96, 64)):
        insts[i]["N1"], insts[i]["N2"] = N, 160 - N
    sw = T.TransmonSweep(contexts=3, k_leaf=8, k_top=12)
    tops, wall, t_leaves = sw.run_batched(insts, mode="mixed")
    plain, _ = sw.run_plain(insts[:2])
    for inst, top in zip(insts, tops):
        spec = O.two_transmons_spec(inst["E_J1"], inst["E_C1"], inst["n_g1"], inst["E_J2"], inst["E_C2"],
                                    inst["n_g2"], inst["g"])
        leaves = [O.Node([(f"varphi_{i}", O.PeriodicGrid(inst[f"N{i}"], 8))], keep=8, real=False) for i in (1, 2)]
        ref = O.materialize(O.Node(leaves, keep=12, real=False), spec)
        assert np.max(np.abs(top.energies - ref.energies) / np.abs(ref.energies)) < 1e-9
        for leaf in top.children:
            assert leaf.stats["max_residual"] <= 1e-8
    for p, b in zip(plain, tops):
        assert np.max(np.abs(p.energies - b.energies) / np.abs(p.energies)) < 1e-9


def _periodic_transmon(N, EJ, EC, ng, ctx=None):
    import kfc
    b = kfc.PeriodicGrid(N, 8)
    return kfc.KronFreeOperator([N], [kfc.AxisTerm(4 * EC, [(0, kfc.localop(b, ("n2", 0)))]),
                                      kfc.AxisTerm(-8 * EC * ng, [(0, kfc.localop(b, ("n", 0)))]),
                                      kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", 1)))]),
                                      kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", -1)))])],
                                constant=4 * EC * ng ** 2, ctx=ctx)


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_chebyshev_batch_complex_vs_dense(mode):
    """Uneven complex periodic grids in one batch: every problem's lowest pairs
    match Dense, with orthonormal vectors and residuals within tolerance."""
    import kfc
    params = [(96, 25.0, 1.0, 0.31), (160, 30.0, 1.2, 0.05), (64, 20.0, 0.9, 0.77), (200, 33.0, 1.3, 0.5)]
    Hs = [_periodic_transmon(*p) for p in params]
    res, st = kfc.solve_chebyshev_batch(Hs, 6, extra=6, degree=40, tol=1e-9, mode=mode)
    for H, (E, V), s in zip(Hs, res, st):
        Ed, _ = kfc.solve(kfc.Dense(), H, 6)
        assert np.max(np.abs(E - Ed)) < 1e-9 * max(1.0, np.max(np.abs(Ed)))
        assert np.allclose(V.conj().T @ V, np.eye(6), atol=1e-9)
        r = kfc.residuals(H, E, V)
        assert np.max(r) < 1e-8 and s["max_residual"] <= 1e-9


def test_chebyshev_batch_real_charge_basis():
    """Real one-axis operators (charge-basis transmons: shifts and a diagonal)."""
    import kfc
    Hs = []
    for EJ, EC, ng in ((30.0, 1.2, 0.0), (12.0, 0.5, 0.0), (40.0, 2.0, 0.0)):
        b = kfc.ChargeBasis(30)
        Hs.append(kfc.KronFreeOperator([b.dim], [kfc.AxisTerm(4 * EC, [(0, kfc.localop(b, ("n2", 0)))]),
                                                 kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", 1)))]),
                                                 kfc.AxisTerm(-EJ / 2, [(0, kfc.localop(b, ("expiphi", -1)))])]))
    assert all(H.dtype == np.float64 for H in Hs)
    res, st = kfc.solve_chebyshev_batch(Hs, 5, extra=5, degree=30, tol=1e-10, mode="f64")
    for H, (E, V), s in zip(Hs, res, st):
        Ed, _ = kfc.solve(kfc.Dense(), H, 5)
        assert np.max(np.abs(E - Ed)) < 1e-10 * max(1.0, np.max(np.abs(Ed)))
        assert np.max(kfc.residuals(H, E, V)) < 1e-9


def test_chebyshev_batch_rejects():
    import kfc
    from kfc import configs
    H1 = _periodic_transmon(64, 25.0, 1.0, 0.3)
    H3 = configs.three_fluxonia_operator(configs.three_fluxonia_theta(1), N=8)
    with pytest.raises(kfc.UnsupportedError):
        kfc.solve_chebyshev_batch([H3], 2, extra=2)
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_chebyshev_batch([H1], 40, extra=40)     # m + extra above N
    with pytest.raises(kfc.UnsupportedError):
        kfc.solve_chebyshev_batch([H1], 20, extra=20)     # more than 32 block columns
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_chebyshev_batch([H1], 4, extra=4, degree=0)


def test_transmon_batch_chebyshev_vs_oracle():
    from oracle import kfc_oracle as O
    from kfc import transmon_sweep as T
    insts = T.draw(3, seed=12)
    for i, N in enumerate((64, 112, 80)):
        insts[i]["N1"], insts[i]["N2"] = N, 176 - N
    sw = T.TransmonSweep(contexts=1, k_leaf=8, k_top=12)
    tops, _, _ = sw.run_batched(insts, mode="mixed", method="chebyshev", degree=40)
    for inst, top in zip(insts, tops):
        spec = O.two_transmons_spec(inst["E_J1"], inst["E_C1"], inst["n_g1"], inst["E_J2"], inst["E_C2"],
                                    inst["n_g2"], inst["g"])
        leaves = [O.Node([(f"varphi_{i}", O.PeriodicGrid(inst[f"N{i}"], 8))], keep=8, real=False) for i in (1, 2)]
        ref = O.materialize(O.Node(leaves, keep=12, real=False), spec)
        assert np.max(np.abs(top.energies - ref.energies) / np.abs(ref.energies)) < 1e-9


def test_learned_subspace_start():
    """The learned-subspace start (kfc_lobpcg_lockstep_subspace): each point begins at
    the lowest Ritz vectors of its operator on span(P B C), B the leading principal
    directions of training eigenvectors in the canonical frame.  Same energies and
    residual bar as the cold lockstep sweep, fewer iterations; a coefficient block
    selecting the first ns columns of B gives the solve of B[:, :ns] bit for bit."""
    import torch

    import kfc
    from kfc import configs
    from kfc.sweeps import LearnedSubspace
    sw = _small_sweep()
    W = sw.width(12 * 12 * 10)
    train = [configs.zero_pi_theta(i) for i in range(48)]
    test = [configs.zero_pi_theta(9100 + i) for i in range(6)]
    gen, _ = sw.run_batched(train, mode="f64", nev=W)
    frames = [sw.to_frame(V, sw.frame_index(kids)) for E, V, st, kids in gen]
    sub = LearnedSubspace(sw, nbasis=48).fit(frames)
    assert sub.B.shape == (frames[0].shape[0], 48)
    B = sub.B.cpu().numpy()
    assert np.allclose(B.T @ B, np.eye(48), atol=1e-12)
    cold, _ = sw.run_batched(test, mode="f64", lockstep=True)
    warm, _ = sw.run_batched(test, mode="f64", subspace=sub)
    for (Ec, Vc, sc, _), (Ew, Vw, sw_, _) in zip(cold, warm):
        assert sc["max_residual"] <= 1e-8 and sw_["max_residual"] <= 1e-8
        assert np.max(np.abs(Ec - Ew) / np.abs(Ec)) < 1e-9
        assert np.allclose(Vw.T @ Vw, np.eye(sw.m), atol=1e-10)
        assert sw_["warm"] == 1
    it_cold = np.mean([o[2]["iterations"] for o in cold])
    it_warm = np.mean([o[2]["iterations"] for o in warm])
    assert it_warm < 0.8 * it_cold, (it_warm, it_cold)
    # C selecting B's first ns columns == the subspace of those columns
    ns = 24
    Hs, kidss = zip(*[sw.parent(th, ctx=sw.ctxs[0]) for th in test])
    maps = [sw.frame_index(k) for k in kidss]
    s = kfc.LOBPCG(tol=1e-8, mode="f64", guard=sw.guard)
    a, sa = kfc.solve_batch_subspace(s, list(Hs), sw.m, sub.B[:, :ns].contiguous(), maps)
    Cs = [np.eye(48)[:, :ns] for _ in test]
    b, sb = kfc.solve_batch_subspace(s, list(Hs), sw.m, sub.B, maps, Cs)
    for (Ea, Va), (Eb, Vb), x, y in zip(a, b, sa, sb):
        assert np.array_equal(Ea, Eb) and np.array_equal(Va, Vb)
        assert x["iterations"] == y["iterations"]
    # refusals: a subspace narrower than the block, two frame rows on one parent row
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch_subspace(s, list(Hs), sw.m, sub.B[:, :W - 1].contiguous(), maps)
    bad = [m_.copy() for m_ in maps]
    hit = np.flatnonzero(bad[0] >= 0)
    bad[0][hit[1]] = bad[0][hit[0]]
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch_subspace(s, list(Hs), sw.m, sub.B, bad)
    with pytest.raises(kfc.DimensionMismatch):
        kfc.solve_batch_subspace(s, list(Hs), sw.m, sub.B, [m_[:-1] for m_ in maps])
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch_subspace(s, list(Hs), sw.m, torch.zeros(4, 4), maps)
