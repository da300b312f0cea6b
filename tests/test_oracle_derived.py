"""Pins the oracle's derived quantities (src/apply.jl:586-685, src/solvers.jl:75)
against brute force, so that the GPU forms compared with them are pinned too:

* ``infnorm`` equals the maximum row sum of ``|dense(op)|`` for every LocalOp kind;
* ``gershgorin_bound`` is an EQUALITY on operators with one kept term
  (||A (x) B||_inf = ||A||_inf ||B||_inf), the sum of the kept terms' row-sum norms
  otherwise, and equal to lambda_max where the bound is attained (c J (x) J);
* ``separable_eigenbases`` diagonalizes the single-axis sums built here from dense
  matrices, with the constant on axis 0 (test/test_apply.jl:487-524);
* ``product_seed`` columns are the product states of the k smallest Kronecker sums
  in ``partialsortperm`` order (sum, then flat index with axis 0 fastest), checked
  against an enumeration of every index tuple, with exact ties between axes;
* the branch form (every basis the identity) takes unit vectors of the k smallest
  entries of the FULL diagonal, kept terms included (test/test_apply.jl:447-461);
* ``residual_floor`` = 10 eps max(|hbound|, |E1|) (src/solvers.jl:75,
  test/test_solvers.jl:99-100).

A flattening in C order, a wrong tie rule, a dropped constant or a bound twice
too loose fails one of these."""
import itertools

import numpy as np
import pytest

from oracle import kfc_oracle as O

EPS = np.finfo(float).eps


def _herm(rng, d, cplx=True):
    A = rng.standard_normal((d, d)) + (1j * rng.standard_normal((d, d)) if cplx else 0)
    return (A + A.conj().T) / 2


def _row_sum_norm(M):
    return float(np.max(np.sum(np.abs(M), axis=1)))


def _ops(rng):
    return [O.DenseOp(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))),
            O.DiagonalOp(rng.standard_normal(5)), O.DiagonalOp(np.array([1, -2, 0], dtype=complex)),
            O.BandedOp(5, [-1, 0, 2], [rng.standard_normal(5) + 1j * rng.standard_normal(5) for _ in range(3)]),
            O.BandedOp(4, [-3, 1], [rng.standard_normal(4) for _ in range(2)]),
            O.ShiftOp(5, -2, 0.5j), O.ShiftOp(3, 1, 0.3), O.ShiftOp(3, 4, 1.0)]


def test_infnorm_every_kind():
    rng = np.random.default_rng(72)
    for op in _ops(rng):
        assert abs(O.infnorm(op) - _row_sum_norm(O.dense(op))) <= 1e-14 * max(1.0, O.infnorm(op))
    # the reference's literals (test/test_apply.jl:462)
    assert O.infnorm(O.ShiftOp(5, 1, 0.3)) == 0.3
    assert O.infnorm(O.DiagonalOp(np.array([1, -2, 0], dtype=complex))) == 2


def _two_axis_fixture(rng, coeff=0.3 + 0.2j):
    dims = [4, 3]
    terms = [O.AxisTerm(1.0, [(0, O.DiagonalOp(rng.standard_normal(4) + 1j * rng.standard_normal(4)))]),
             O.AxisTerm(0.5, [(1, O.DiagonalOp(rng.standard_normal(3)))]),
             O.AxisTerm(coeff, [(0, O.DenseOp(rng.standard_normal((4, 4)))),
                                (1, O.BandedOp(3, [-1, 1], [rng.standard_normal(3), rng.standard_normal(3)]))])]
    return dims, terms


def test_gershgorin_single_kept_term_is_an_equality():
    rng = np.random.default_rng(5)
    for trial in range(4):
        dims, terms = _two_axis_fixture(rng, coeff=complex(*rng.standard_normal(2)))
        H = O.kron_operator(dims, terms, constant=0.25 - 0.5j)
        assert len(H.terms) == 1
        kept = O.kron_reference(dims, H.terms)            # c A (x) B, dense
        fused = np.diag(O.kron_reference(dims, H.diagterms))   # the constant is a fused term
        expect = np.max(fused.real) + np.max(np.abs(fused.imag)) + _row_sum_norm(kept)
        assert abs(O.gershgorin_bound(H) - expect) <= 1e-13 * abs(expect)


def test_gershgorin_sum_over_terms_and_attained():
    rng = np.random.default_rng(6)
    dims = [3, 4, 2]
    terms = [O.AxisTerm(0.7, [(0, O.DenseOp(_herm(rng, 3)))]),
             O.AxisTerm(-0.2, [(1, O.BandedOp(4, [-1, 0, 1], [rng.standard_normal(4) for _ in range(3)])),
                               (2, O.ShiftOp(2, 1, 0.5))]),
             O.AxisTerm(1.1j, [(0, O.ShiftOp(3, -1, 2.0)), (2, O.DenseOp(rng.standard_normal((2, 2))))])]
    H = O.kron_operator(dims, terms, constant=1.5)
    off = sum(_row_sum_norm(O.kron_reference(dims, [t])) for t in H.terms)
    assert abs(O.gershgorin_bound(H) - (1.5 + off)) <= 1e-13 * (1.5 + off)
    # c J (x) J on [d1, d2]: lambda_max = c d1 d2 = the row sum, so the bound is attained
    for d1, d2, c in ((3, 4, 0.5), (5, 2, 2.0)):
        Hj = O.kron_operator([d1, d2], [O.AxisTerm(c, [(0, O.DenseOp(np.ones((d1, d1)))),
                                                        (1, O.DenseOp(np.ones((d2, d2))))])])
        lmax = np.max(np.linalg.eigvalsh(O.kron_reference([d1, d2], Hj.terms)))
        assert abs(O.gershgorin_bound(Hj) - lmax) <= 1e-12 * lmax
        assert abs(lmax - c * d1 * d2) <= 1e-12 * lmax


def _reference_fixture(rng):
    """The operator of test/test_apply.jl:489-505: non-diagonal single-axis terms on
    every axis, a fused two-axis diagonal coupling, a banded (x) shift coupling with
    its adjoint, and a constant."""
    dims = [4, 5, 3]
    A1 = _herm(rng, 4)
    w = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    B2 = O.BandedOp(5, [-1, 0, 1], [np.concatenate([[0], np.conj(w[:4])]), rng.standard_normal(5) + 0j,
                                    np.concatenate([w[:4], [0]])])
    Bc = O.BandedOp(4, [0, 2], [rng.standard_normal(4) + 1j * rng.standard_normal(4) for _ in range(2)])
    Bc_adj = O.DenseOp(O.dense(Bc).conj().T)
    c = 0.3 + 0.2j
    S1, Sm1 = O.ShiftOp(3, 1, 1.0), O.ShiftOp(3, -1, 1.0)
    terms = [O.AxisTerm(1.0, [(0, O.DenseOp(A1))]), O.AxisTerm(0.5, [(1, B2)]),
             O.AxisTerm(0.4, [(2, S1)]), O.AxisTerm(0.4, [(2, Sm1)]),
             O.AxisTerm(0.2, [(0, O.DiagonalOp(rng.standard_normal(4))), (1, O.DiagonalOp(rng.standard_normal(5)))]),
             O.AxisTerm(c, [(0, Bc), (2, S1)]), O.AxisTerm(np.conj(c), [(0, Bc_adj), (2, Sm1)])]
    single = [A1 + 0.7 * np.eye(4), 0.5 * O.dense(B2), 0.4 * (O.dense(S1) + O.dense(Sm1))]
    return dims, terms, single


def test_separable_eigenbases_and_seed_reference_fixture():
    rng = np.random.default_rng(72)
    dims, terms, single = _reference_fixture(rng)
    H = O.kron_operator(dims, terms, constant=0.7)
    ref = O.kron_reference(dims, terms, constant=0.7)
    assert np.linalg.norm(ref - ref.conj().T) <= 1e-14 * np.linalg.norm(ref)
    assert len(H.diagterms) == 2            # the fused coupling and the constant
    assert O.gershgorin_bound(H) >= np.max(np.linalg.eigvalsh(ref))
    U, lam = O.separable_eigenbases(H)
    for k in range(3):
        assert np.allclose(U[k].conj().T @ U[k], np.eye(dims[k]), atol=1e-13)
        assert np.allclose(U[k] @ np.diag(lam[k]) @ U[k].conj().T, single[k], atol=1e-13)
        assert not np.allclose(U[k], np.eye(dims[k]))
    Hsep = O.kron_reference(dims, [O.AxisTerm(1.0, [(k, O.DenseOp(single[k]))]) for k in range(3)])
    sums = sorted(a + b + d for a in np.linalg.eigvalsh(single[0]) for b in np.linalg.eigvalsh(single[1])
                  for d in np.linalg.eigvalsh(single[2]))
    X = O.product_seed(H, 5)
    assert X.shape == (60, 5) and np.allclose(X.conj().T @ X, np.eye(5), atol=1e-13)
    P = X.conj().T @ Hsep @ X
    assert np.linalg.norm(P - np.diag(sums[:5])) <= 1e-12


def test_separable_refuses_non_hermitian():
    Hn = O.kron_operator([3], [O.AxisTerm(1.0, [(0, O.ShiftOp(3, 1, 1.0))])])
    with pytest.raises(ValueError, match="Hermitian"):
        O.separable_eigenbases(Hn)


def _rank1_factors(v, dims):
    """The per-axis factors of a product state v (axis 0 fastest), each unit norm."""
    T = v.reshape(dims, order="F")
    out = []
    for a in range(len(dims)):
        M = np.moveaxis(T, a, 0).reshape(dims[a], -1)
        u, s, _ = np.linalg.svd(M)
        assert s[1:].sum() <= 1e-10 * s[0] if len(s) > 1 else True, "seed column is not a product state"
        out.append(u[:, 0])
    return out


def _spectrum_with_gaps(rng, d, lo=0.0):
    """d eigenvalues at least 0.05 apart (distinct sums of distinct tuples differ by
    far more than rounding unless they tie exactly)."""
    return lo + np.cumsum(0.05 + rng.uniform(0, 1, d))


@pytest.mark.parametrize("dims", [[5, 5], [4, 4, 3], [3, 3, 3]])
def test_product_seed_order_against_enumeration(dims):
    """Axes 0 and 1 carry the same single-axis matrix, so lambda_0[i] + lambda_1[j]
    and lambda_0[j] + lambda_1[i] tie exactly; partialsortperm keeps the lower flat
    index (axis 0 fastest) first.  Each seed column must be the product state of the
    tuple the enumeration puts at its position, up to a phase per factor."""
    rng = np.random.default_rng(sum(dims))
    mats = []
    for a, d in enumerate(dims):
        if a == 1 and dims[1] == dims[0]:
            mats.append(mats[0])
            continue
        Q, _ = np.linalg.qr(rng.standard_normal((d, d)))
        mats.append(Q @ np.diag(_spectrum_with_gaps(rng, d)) @ Q.T)
    terms = [O.AxisTerm(1.0, [(a, O.DenseOp(M))]) for a, M in enumerate(mats)]
    # a coupling term must not change the seed (separable part only)
    terms.append(O.AxisTerm(0.3, [(0, O.DenseOp(_herm(rng, dims[0], False))),
                                  (len(dims) - 1, O.DenseOp(_herm(rng, dims[-1], False)))]))
    H = O.kron_operator(dims, terms)
    ev = [np.linalg.eigvalsh(M) for M in mats]
    strides = np.cumprod([1] + dims[:-1])
    tuples = list(itertools.product(*[range(d) for d in dims]))
    key = {t: (round(float(sum(ev[a][t[a]] for a in range(len(dims)))), 9), int(np.dot(t, strides))) for t in tuples}
    order = sorted(tuples, key=lambda t: key[t])
    ties = sum(1 for x, y in zip(order, order[1:]) if key[x][0] == key[y][0])
    assert ties > 0 or dims[0] != dims[1]
    k = min(12, len(tuples))
    X = O.product_seed(H, k)
    for j in range(k):
        facs = _rank1_factors(X[:, j], dims)
        for a, u in enumerate(facs):
            r = float(np.real(u.conj() @ mats[a] @ u))
            assert np.linalg.norm(mats[a] @ u - r * u) <= 1e-10, "seed factor is not an eigenvector"
            i_a = int(np.argmin(np.abs(ev[a] - r)))
            assert i_a == order[j][a], f"column {j}: axis {a} index {i_a}, enumeration says {order[j]}"


def test_product_seed_branch_unit_vectors():
    """test/test_apply.jl:447-461: a branch (dense two-axis factors only) seeds with
    unit vectors of the smallest diagonal entries; a kept zero-shift ShiftOp (x)
    DiagonalOp term, not fused into the stored diagonal, reorders them."""
    rng = np.random.default_rng(71)
    termsb = [O.AxisTerm(1.0, [(0, O.DiagonalOp(np.arange(1, 5).astype(complex)))]),
              O.AxisTerm(1.0, [(1, O.DiagonalOp(np.array([0.0, 5.0, 10.0], dtype=complex)))]),
              O.AxisTerm(0.1, [(0, O.DenseOp(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))),
                               (1, O.DenseOp(rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))))])]
    Hb = O.kron_operator([4, 3], termsb)
    Xb = O.product_seed(Hb, 3)
    assert all(np.count_nonzero(Xb[:, j]) == 1 for j in range(3))
    assert int(np.flatnonzero(Xb[:, 0])[0]) == int(np.argmin(np.real(Hb.diag)))
    Hk = O.kron_operator([4, 3], termsb + [O.AxisTerm(1.0, [(0, O.ShiftOp(4, 0, 1.0)),
                                                             (1, O.DiagonalOp(np.array([0.0, -20.0, 0.0])))])])
    Dk = np.real(np.diag(O.kron_reference([4, 3], termsb + [O.AxisTerm(1.0, [(0, O.ShiftOp(4, 0, 1.0)),
                                                                                (1, O.DiagonalOp([0.0, -20.0, 0.0]))])])))
    assert np.allclose(O.full_diagonal(Hk), Dk)
    assert int(np.argmin(Dk)) != int(np.argmin(np.real(Hk.diag)))
    Xk = O.product_seed(Hk, 3)
    assert [int(np.flatnonzero(Xk[:, j])[0]) for j in range(3)] == list(np.argsort(Dk, kind="stable")[:3])


def test_product_seed_branch_ties_keep_index_order():
    """Integer diagonals with repeated values: ties between unit vectors go to the
    lower flat index, as partialsortperm's stable order does."""
    rng = np.random.default_rng(3)
    d0 = np.array([2.0, 0.0, 1.0, 0.0])
    d1 = np.array([1.0, 0.0, 1.0])
    terms = [O.AxisTerm(1.0, [(0, O.DiagonalOp(d0))]), O.AxisTerm(1.0, [(1, O.DiagonalOp(d1))]),
             O.AxisTerm(0.1, [(0, O.DenseOp(_herm(rng, 4, False) - np.diag(np.diag(_herm(rng, 4, False))))),
                              (1, O.DenseOp(np.array([[0.0, 1, 0], [1, 0, 1], [0, 1, 0]])))])]
    H = O.kron_operator([4, 3], terms)
    D = np.real(np.diag(O.kron_reference([4, 3], terms)))
    expect = sorted(range(12), key=lambda i: (D[i], i))[:7]
    X = O.product_seed(H, 7)
    assert [int(np.flatnonzero(X[:, j])[0]) for j in range(7)] == expect


def test_residual_floor_closed_form():
    # test/test_solvers.jl:99-100: 10 eps ||H|| with ||H|| the Gershgorin bound when it dominates
    H = O.three_fluxonia_operator([1.0, 0.9, 1.1], [0.5, 0.4, 0.6], [4, 5, 6], [0.1, 0.2, 0.3], N=10, order=2)
    hb = O.gershgorin_bound(H)
    E1 = float(np.linalg.eigvalsh(O.to_matrix(H))[0])
    assert abs(E1) < abs(hb)
    assert O.residual_floor(hb, E1) == 10 * EPS * hb
    # |E1| dominates: every state of c I (c < 0) has E = c = the bound; a deep well below a
    # negative bound takes |E1|
    assert O.residual_floor(-5.0, -5.0) == 10 * EPS * 5.0
    Hw = O.kron_operator([4], [O.AxisTerm(1.0, [(0, O.DiagonalOp(np.array([-100.0, -3.0, -2.0, -1.0])))]),
                               O.AxisTerm(0.1, [(0, O.BandedOp(4, [-1, 1], [np.ones(4), np.ones(4)]))])])
    hbw = O.gershgorin_bound(Hw)
    E1w = float(np.linalg.eigvalsh(O.to_matrix(Hw))[0])
    assert hbw < 0 and abs(E1w) > abs(hbw)
    assert O.residual_floor(hbw, E1w) == 10 * EPS * abs(E1w)
    assert O.residual_floor(3.0, -7.0) == O.residual_floor(-3.0, 7.0) == 10 * EPS * 7.0
