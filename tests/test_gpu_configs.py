"""The workloads as configured, against the oracle's own specs and materialize:

* the batched transmon sweep at its grid sizes ({256, 384, 512, 640, 768}, 768
  included) through kfc_chebyshev_batch leaves and Dense tops keeping 20: every top
  equals the oracle's dense tree to 1e-9 relative;
* the product's spec builders (kfc.configs) against the oracle's specs of the
  reference fixtures (test/fixtures/full_zero_pi.json, two_transmons.json) and the
  three-fluxonia form: same numbers, same materialized energies;
* device ports of a zero-pi parent (kfc_sandwich) against the oracle's port_matrix /
  joint_port (src/materialize.jl:443-518) on the oracle's own tree, each kept state's
  phase fixed by lifting both sets of parent states to the leaves' primitive basis."""
import math

import numpy as np
import pytest

from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu


def test_transmon_sweep_at_configured_grid_sizes():
    from kfc import transmon_sweep as T
    insts = T.draw(7, seed=4242)
    grids = [(768, 256), (640, 768), (512, 384), (384, 640), (256, 512), (768, 768), (256, 256)]
    for inst, (n1, n2) in zip(insts, grids):
        inst["N1"], inst["N2"] = n1, n2
    sw = T.TransmonSweep(contexts=1)
    tops, wall, t_leaves = sw.run_batched(insts, mode="mixed", method="chebyshev")
    assert len(tops) == len(insts)
    for inst, top in zip(insts, tops):
        spec = O.two_transmons_spec(inst["E_J1"], inst["E_C1"], inst["n_g1"], inst["E_J2"], inst["E_C2"],
                                    inst["n_g2"], inst["g"])
        leaves = [O.Node([(f"varphi_{i}", O.PeriodicGrid(inst[f"N{i}"], 8))], keep=8, real=False) for i in (1, 2)]
        ref = O.materialize(O.Node(leaves, keep=20, real=False), spec)
        assert len(top.energies) == 20
        assert np.max(np.abs(top.energies - ref.energies) / np.abs(ref.energies)) < 1e-9, (inst["N1"], inst["N2"])
        for leaf, rleaf in zip(top.children, ref.children):
            assert leaf.stats["max_residual"] <= 1e-8
            assert np.max(np.abs(leaf.energies - rleaf.energies) / np.abs(rleaf.energies)) < 1e-9


def test_zero_pi_spec_builder_matches_oracle_fixture_spec():
    import kfc
    from kfc import configs
    for seed in (0, 17):
        th = configs.zero_pi_theta(seed)
        spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
        ospec = O.full_zero_pi_spec(E_J=10.0, E_L=0.05, E_CJ=20.0, E_CS=th["E_CS"], E_Cz=E_Cz, dE_J=0.05, dC_J=0.0,
                                    dC=0.0, dE_L=0.1, varphi_ext=2 * math.pi * 0.23, n_g=th["n_g"])
        assert abs(E_Cz - 1 / (1 / th["E_CS"] - 1 / 20.0)) < 1e-15
        assert np.allclose(spec.E_C, ospec.E_C, rtol=0, atol=1e-15)
        assert np.allclose(spec.E_L, ospec.E_L, rtol=0, atol=1e-15)
        assert np.allclose(spec.n_g, ospec.n_g, rtol=0, atol=1e-15)
        assert abs(spec.constant - ospec.constant) < 1e-12
        assert [(float(e), list(w), float(o)) for e, w, o in spec.cosines] == \
               [(float(e), list(w), float(o)) for e, w, o in ospec.cosines]
        # both materialize to the same energies on a small tree (device LOBPCG parent vs oracle Dense)
        root = configs.zero_pi_tree(E_Cz, ncut=8, Nphi=48, Nzeta=8, keep=8, m=6,
                                    solver=kfc.LOBPCG(tol=1e-9))
        s = kfc.materialize(root, spec)
        oroot = O.Node([O.Node([("theta", O.ChargeBasis(8))], keep=8),
                        O.Node([("phi", O.PhaseGrid(-6 * math.pi, 6 * math.pi, 48, 8))], keep=8),
                        O.Node([("zeta", O.FockBasis(8, (2 * E_Cz / 0.05) ** 0.25))], keep=8)], keep=6)
        ref = O.materialize(oroot, ospec)
        assert np.max(np.abs(s.energies() - ref.energies) / np.abs(ref.energies)) < 1e-9


def test_two_transmons_and_three_fluxonia_spec_builders():
    import kfc
    from kfc import configs
    th = configs.TWO_TRANSMONS_THETA
    a, b = configs.two_transmons_spec(**th), O.two_transmons_spec(**th)
    assert np.array_equal(np.asarray(a.E_C), np.asarray(b.E_C)) and np.array_equal(a.n_g, b.n_g)
    t3 = configs.three_fluxonia_theta(20261203)
    a = configs.three_fluxonia_spec(t3["E_C"], t3["E_L"], t3["E_J"], t3["phi_ext"])
    b = O.three_fluxonia_spec(t3["E_C"], t3["E_L"], t3["E_J"], t3["phi_ext"])
    assert np.array_equal(np.asarray(a.E_C), np.asarray(b.E_C)) and np.array_equal(np.asarray(a.E_L), np.asarray(b.E_L))
    assert [(float(e), list(w), float(o)) for e, w, o in a.cosines] == \
           [(float(e), list(w), float(o)) for e, w, o in b.cosines]
    # the device operator of the configured node against the oracle's materialize of its own spec
    n = 12
    H = configs.three_fluxonia_operator(t3, N=n)
    E, V = kfc.solve(kfc.Lanczos(tol=1e-10), H, 6)
    onode = O.Node([(f"phi{i}", O.PhaseGrid(-6 * math.pi, 6 * math.pi, n, 2)) for i in (1, 2, 3)], keep=6)
    ref = O.materialize(onode, b)
    assert np.max(np.abs(E - ref.energies) / np.abs(ref.energies)) < 1e-9


def _lift(mat):
    """The kept states of a materialized node in its leaves' primitive product basis
    (axis 0 fastest): each node child's kept vectors expanded, recursively."""
    V = np.asarray(mat.vectors)
    facs = []
    for ch, kid in zip(mat.node.children, mat.children):
        if isinstance(ch, tuple):
            facs.append(np.eye(ch[1].dim))
        else:
            facs.append(_lift(kid))
    K = np.ones((1, 1))
    for F in facs:
        K = np.kron(F, K)
    return K @ V


@pytest.fixture(scope="module")
def zero_pi_pair():
    import kfc
    from kfc import configs
    th = configs.zero_pi_theta(8800)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    # a Dense parent (512 states) leaves the kept states exact to rounding; the kept gap
    # is 0.0106, so an iterative solve at tol 1e-10 would move them by ~1e-8
    root = configs.zero_pi_tree(E_Cz, ncut=8, Nphi=48, Nzeta=8, keep=8, m=6, solver=kfc.Dense())
    s = kfc.materialize(root, spec).root
    ospec = O.full_zero_pi_spec(E_J=10.0, E_L=0.05, E_CJ=20.0, E_CS=th["E_CS"], E_Cz=E_Cz, dE_J=0.05, dC_J=0.0,
                                dC=0.0, dE_L=0.1, varphi_ext=2 * math.pi * 0.23, n_g=th["n_g"])
    oroot = O.Node([O.Node([("theta", O.ChargeBasis(8))], keep=8),
                    O.Node([("phi", O.PhaseGrid(-6 * math.pi, 6 * math.pi, 48, 8))], keep=8),
                    O.Node([("zeta", O.FockBasis(8, (2 * E_Cz / 0.05) ** 0.25))], keep=8)], keep=6)
    ref = O.materialize(oroot, ospec)
    Uk, Uo = _lift(s), _lift(ref)
    S = Uo.conj().T @ Uk                    # the per-state phases (diagonal, unit modulus)
    assert np.allclose(np.abs(np.diag(S)), 1.0, atol=1e-10)
    assert np.linalg.norm(S - np.diag(np.diag(S))) < 1e-10
    return s, ref, np.diag(np.diag(S))


@pytest.mark.parametrize("mode,port", [("theta", ("n", 0)), ("theta", ("expiphi", 1)), ("phi", ("phi", 1)),
                                       ("phi", ("n", 0)), ("zeta", ("n", 0)), ("zeta", ("phi2", 0))])
def test_parent_ports_device_vs_oracle(zero_pi_pair, mode, port):
    from kfc.tree import port_matrix
    s, ref, S = zero_pi_pair
    dev = port_matrix(s, mode, port, device=True)
    want = S.conj().T @ O.port_matrix(ref, mode, port) @ S
    scale = max(1.0, float(np.max(np.abs(want))))
    assert np.max(np.abs(dev - want)) <= 1e-10 * scale


def test_parent_joint_ports_device_vs_oracle(zero_pi_pair):
    from kfc.tree import joint_port
    s, ref, S = zero_pi_pair
    for factors in ([("theta", ("expiphi", 1)), ("phi", ("expiphi", 1))],
                    [("theta", ("n", 0)), ("zeta", ("n", 0))],
                    [("theta", ("n", 0)), ("phi", ("phi", 1)), ("zeta", ("phi", 1))]):
        dev = joint_port(s, factors, device=True)
        want = S.conj().T @ O.joint_port(ref, factors) @ S
        scale = max(1.0, float(np.max(np.abs(want))))
        assert np.max(np.abs(dev - want)) <= 1e-10 * scale
