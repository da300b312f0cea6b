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
