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
