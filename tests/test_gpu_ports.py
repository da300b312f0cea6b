"""Port formation on the device (kfc_sandwich): V' (x_a O_a) V of a materialized
node, the quantity _port_matrix / _joint_port form on the host
(src/materialize.jl:441-518).  Every device port equals the host formula (which
tests/test_oracle.py pins against scqubits through the tree energies) to
rounding, for single and joint ports, real and complex operators on real nodes,
complex nodes, and node children (a parent's ports go through its children's)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _close(A, B, rtol=1e-12):
    A, B = np.asarray(A), np.asarray(B)
    assert A.shape == B.shape
    scale = max(1.0, float(np.max(np.abs(B))))
    assert np.max(np.abs(A - B)) <= rtol * scale, float(np.max(np.abs(A - B)))


@pytest.fixture(scope="module")
def zero_pi_parent():
    """A zero-pi parent [40, 40, 30] materialized (48 000 states, 16 kept)."""
    import kfc
    from kfc import configs
    th = configs.zero_pi_theta(8800)
    spec, E_Cz = configs.zero_pi_spec(th["E_CS"], th["n_g"])
    return kfc.materialize(configs.zero_pi_tree(E_Cz), spec).root


@pytest.mark.parametrize("mode,port", [("theta", ("n", 0)), ("theta", ("expiphi", 1)), ("phi", ("phi", 1)),
                                       ("phi", ("n", 0)), ("zeta", ("n", 0)), ("zeta", ("phi2", 0))])
def test_parent_ports_device_vs_host(zero_pi_parent, mode, port):
    from kfc.tree import port_matrix
    mn = zero_pi_parent
    assert mn.vectors.shape == (48000, 16)
    dev = port_matrix(mn, mode, port, device=True)
    host = port_matrix(mn, mode, port, device=False)
    _close(dev, host)
    if port[0] in ("n", "phi", "phi2"):   # Hermitian ports stay Hermitian
        _close(dev, np.asarray(dev).conj().T)


def test_parent_joint_port_device_vs_host(zero_pi_parent):
    from kfc.tree import joint_port
    mn = zero_pi_parent
    for factors in ([("theta", ("expiphi", 1)), ("phi", ("expiphi", 1))],
                    [("theta", ("n", 0)), ("zeta", ("n", 0))],
                    [("theta", ("n", 0)), ("phi", ("phi", 1)), ("zeta", ("phi", 1))]):
        _close(joint_port(mn, factors, device=True), joint_port(mn, factors, device=False))


def test_complex_node_ports():
    """A complex node (coupled transmons on periodic grids: n is imaginary there)."""
    import kfc
    from kfc import transmon_sweep as T
    from kfc.tree import joint_port, port_matrix
    inst = T.draw(1, seed=5)[0]
    inst["N1"], inst["N2"] = 256, 384
    top = kfc.materialize(T.tree_of(inst, k_leaf=8, k_top=20), T.spec_of(inst)).root
    assert np.iscomplexobj(top.vectors)
    for mode in ("varphi_1", "varphi_2"):
        for port in (("n", 0), ("expiphi", 1)):
            _close(port_matrix(top, mode, port, device=True), port_matrix(top, mode, port, device=False))
    f = [("varphi_1", ("expiphi", 1)), ("varphi_2", ("expiphi", -1))]
    _close(joint_port(top, f, device=True), joint_port(top, f, device=False))
    # and on a leaf (a basis child), its own vectors
    leaf = top.children[0]
    _close(port_matrix(leaf, "varphi_1", ("n", 0), device=True), port_matrix(leaf, "varphi_1", ("n", 0), device=False))


def test_sandwich_rejects():
    import kfc
    from kfc import configs
    from kfc import _lib as L
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(1), N=8)
    import ctypes as C
    import torch
    V = torch.zeros(3, H.N, dtype=torch.float64, device="cuda")
    G = np.empty(9)
    with pytest.raises(kfc.ArgumentError):
        L.check(L.lib().kfc_sandwich(H.handle, C.c_void_p(V.data_ptr()), 0, H.N, G.ctypes.data_as(L._D)),
                H.ctx.handle)
    with pytest.raises(kfc.ArgumentError):
        L.check(L.lib().kfc_sandwich(H.handle, C.c_void_p(V.data_ptr()), 3, H.N - 1, G.ctypes.data_as(L._D)),
                H.ctx.handle)
