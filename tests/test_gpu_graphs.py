"""The batched ChebyshevFilter's filter steps run as one captured hipGraph per
iteration (KFC_GRAPHS unset), replayed while the active set and degree stay the
same, with the eager orthonormalization and Rayleigh-Ritz (and their
synchronizations) between the replays.  The graph path must give exactly the
separate launches' result (KFC_GRAPHS=0): same kernels, same arguments, same
coefficient table.  Both are checked against the oracle's dense leaves."""
import numpy as np
import pytest

from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu


def _run(monkeypatch, graphs, insts, mode):
    from kfc import transmon_sweep as T
    monkeypatch.setenv("KFC_GRAPHS", "1" if graphs else "0")
    sw = T.TransmonSweep(contexts=1, k_leaf=8, k_top=12)
    tops, wall, t_leaves = sw.run_batched(insts, mode=mode, method="chebyshev")
    return tops


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_filter_graph_equals_launches(monkeypatch, mode):
    from kfc import transmon_sweep as T
    insts = T.draw(5, seed=31)
    for i, (n1, n2) in enumerate(((64, 96), (128, 64), (96, 160), (256, 128), (64, 64))):
        insts[i]["N1"], insts[i]["N2"] = n1, n2
    eager = _run(monkeypatch, False, insts, mode)
    graph = _run(monkeypatch, True, insts, mode)
    again = _run(monkeypatch, True, insts, mode)
    for te, tg, ta in zip(eager, graph, again):
        for le, lg, la in zip(te.children, tg.children, ta.children):
            assert np.array_equal(le.energies, lg.energies) and np.array_equal(le.vectors, lg.vectors)
            assert np.array_equal(lg.energies, la.energies) and np.array_equal(lg.vectors, la.vectors)
        assert np.array_equal(te.energies, tg.energies)
    for inst, top in zip(insts, graph):
        spec = O.two_transmons_spec(inst["E_J1"], inst["E_C1"], inst["n_g1"], inst["E_J2"], inst["E_C2"],
                                    inst["n_g2"], inst["g"])
        leaves = [O.Node([(f"varphi_{i}", O.PeriodicGrid(inst[f"N{i}"], 8))], keep=8, real=False) for i in (1, 2)]
        ref = O.materialize(O.Node(leaves, keep=12, real=False), spec)
        assert np.max(np.abs(top.energies - ref.energies) / np.abs(ref.energies)) < 1e-9


def _lockstep(monkeypatch, graphs, pts, mode, sub=None, sw=None):
    monkeypatch.setenv("KFC_GRAPHS", "1" if graphs else "0")
    return sw.run_batched(pts, mode=mode, lockstep=True, subspace=sub)[0]


@pytest.mark.parametrize("mode", ["f64", "mixed"])
def test_lockstep_step_graph_equals_launches(monkeypatch, mode):
    """The lockstep LOBPCG step (uploads, residual, batched dense-pair apply, Grams,
    sums, records download) as one graph per active count, replayed with the host's
    Rayleigh-Ritz, rotation and synchronization between replays: bit for bit the
    separate launches, and the oracle's dense spectrum."""
    from kfc import configs
    from kfc.sweeps import ZeroPiSweep
    from tests.helpers import oracle_from_kfc_terms
    sw = ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, contexts=1)
    pts = [configs.zero_pi_theta(7300 + i) for i in range(7)]
    eager = _lockstep(monkeypatch, False, pts, mode, sw=sw)
    graph = _lockstep(monkeypatch, True, pts, mode, sw=sw)
    again = _lockstep(monkeypatch, True, pts, mode, sw=sw)
    for e, g, a in zip(eager, graph, again):
        assert np.array_equal(e[0], g[0]) and np.array_equal(e[1], g[1])
        assert np.array_equal(g[0], a[0]) and np.array_equal(g[1], a[1])
        assert e[2]["iterations"] == g[2]["iterations"] == a[2]["iterations"]
        assert e[2]["n_matvec"] == g[2]["n_matvec"]
        assert g[2]["max_residual"] <= 1e-8
    assert eager[0][2]["seconds_apply_gpu"] > 0 and graph[0][2]["seconds_apply_gpu"] == -1.0
    for th, (E, V, st, kids) in zip(pts[:3], graph[:3]):
        H, _ = sw.parent(th)
        w = np.linalg.eigvalsh(O.to_matrix(oracle_from_kfc_terms(H)))[:sw.m]
        assert np.max(np.abs(E - w) / np.abs(w)) < 1e-9


def test_lockstep_subspace_graph_equals_launches(monkeypatch):
    from kfc import configs
    from kfc.sweeps import LearnedSubspace, ZeroPiSweep
    sw = ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, contexts=2)
    W = sw.width(12 * 12 * 10)
    gen, _ = sw.run_batched([configs.zero_pi_theta(i) for i in range(32)], mode="f64", nev=W)
    sub = LearnedSubspace(sw, nbasis=40).fit([sw.to_frame(V, sw.frame_index(k)) for E, V, st, k in gen])
    pts = [configs.zero_pi_theta(7400 + i) for i in range(5)]
    eager = _lockstep(monkeypatch, False, pts, "f64", sub=sub, sw=sw)
    graph = _lockstep(monkeypatch, True, pts, "f64", sub=sub, sw=sw)
    for e, g in zip(eager, graph):
        assert np.array_equal(e[0], g[0]) and np.array_equal(e[1], g[1])
        assert e[2]["iterations"] == g[2]["iterations"]


def test_filter_launch_shapes_equal(monkeypatch):
    """The batched filter step's launch shapes (a block per (column, problem) of 256
    or 128 threads, or a persistent grid over the size-ordered work list), pinned one
    at a time and left to the tuner: the same leaves bit for bit, on an uneven batch."""
    from kfc import transmon_sweep as T
    insts = T.draw(6, seed=41)
    for i, (n1, n2) in enumerate(((64, 256), (320, 64), (96, 160), (256, 128), (64, 64), (512, 96))):
        insts[i]["N1"], insts[i]["N2"] = n1, n2
    runs = []
    for shape in ("0", "1", "2", None):
        if shape is None:
            monkeypatch.delenv("KFC_BATCH_SHAPE", raising=False)
        else:
            monkeypatch.setenv("KFC_BATCH_SHAPE", shape)
        runs.append(_run(monkeypatch, True, insts, "mixed"))
    for other in runs[1:]:
        for ta, tb in zip(runs[0], other):
            for la, lb in zip(ta.children, tb.children):
                assert np.array_equal(la.energies, lb.energies) and np.array_equal(la.vectors, lb.vectors)
            assert np.array_equal(ta.energies, tb.energies)
