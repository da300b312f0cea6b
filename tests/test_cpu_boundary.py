"""CPU-side checks of the C ABI boundary: the library loads, exports every
symbol include/kfc.h declares, and the host mirror validates its arguments
like the reference does."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    import kfc
    from kfc import _lib
    hdr = open(os.path.join(ROOT, "include", "kfc.h")).read()
    declared = set(re.findall(r"\b(kfc_[a-z0-9_]+)\s*\(", hdr))
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    assert kfc.missing_symbols() == []


def test_solver_keyword_validation():
    import kfc
    with pytest.raises(kfc.ArgumentError):
        kfc.LOBPCG(tol=-1)
    with pytest.raises(kfc.ArgumentError):
        kfc.LOBPCG(preconditioner="foo")
    with pytest.raises(kfc.ArgumentError):
        kfc.ChebyshevFilter(extra=1)
    with pytest.raises(kfc.ArgumentError):
        kfc.ChebyshevFilter(degree=201)
    with pytest.raises(kfc.ArgumentError):
        kfc.AxisTerm(1.0, [(1, kfc.DiagonalOp([1.0])), (0, kfc.DiagonalOp([1.0]))])
    s = kfc.resolve(kfc.Auto(), {"isleaf": False, "dims": [40, 40, 30], "m": 16})
    assert isinstance(s, kfc.LOBPCG)
    assert isinstance(kfc.resolve(kfc.Auto(), {"isleaf": True, "dims": [61], "m": 8}), kfc.Dense)
    assert isinstance(kfc.resolve(kfc.Auto(), {"isleaf": True, "dims": [4000], "m": 8}), kfc.ChebyshevFilter)


def test_product_bases_match_oracle():
    import kfc
    from oracle import kfc_oracle as O
    cases = [(kfc.ChargeBasis(5), O.ChargeBasis(5), [("n", 0), ("n2", 0), ("expiphi", 1), ("expiphi", -2)]),
             (kfc.PhaseGrid(-3, 3, 21, 4), O.PhaseGrid(-3, 3, 21, 4), [("n", 0), ("n2", 0), ("phi", 0), ("expiphi", 1)]),
             (kfc.PeriodicGrid(16, 8), O.PeriodicGrid(16, 8), [("n", 0), ("n2", 0), ("expiphi", 2)]),
             (kfc.FockBasis(12, 0.8), O.FockBasis(12, 0.8), [("n", 0), ("n2", 0), ("phi", 0), ("phi2", 0)])]
    for b, bo, ports in cases:
        for p in ports:
            assert np.allclose(kfc.dense(kfc.localop(b, p)), O.dense(O.localop(bo, p)), atol=1e-13), (b, p)


def test_real_node_terms_match_oracle_operator():
    """The product's merged real expansion (src/materialize.jl:616-647) gives the
    same operator as the oracle's unmerged one."""
    import kfc
    from kfc import configs
    from oracle import kfc_oracle as O
    from tests.helpers import oracle_from_kfc_terms
    th = configs.three_fluxonia_theta(4)
    spec = configs.three_fluxonia_spec(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"])
    node = configs.three_fluxonia_node(N=7)
    ts = kfc.tree.node_axisterms(node, [None] * 3, kfc.circuit_terms(spec), real=True)

    class _H:
        dims, terms, constant = [7, 7, 7], ts, 0.0
    Hp = oracle_from_kfc_terms(_H)
    Ho = O.three_fluxonia_operator(th["E_C"], th["E_L"], th["E_J"], th["phi_ext"], N=7, order=2)
    assert np.allclose(O.to_matrix(Hp), O.to_matrix(Ho), atol=1e-11)
    assert len(Hp.terms) == 6
