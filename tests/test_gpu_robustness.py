"""The reference's device robustness cases (test/cuda/test_device.jl:163-207, 272-366)
on the device solvers: kfc_lobpcg, kfc_lobpcg_lockstep, kfc_chebyshev, kfc_lanczos.

* S1: exact and wrong-subspace warm starts (the lowest unit vectors, excited ones,
  the solver's own answer) still give the lowest pairs;
* S3: H = 0 and H = c I give valid pairs or the reference's informative rejection
  (an error naming Dense); tight clusters delta in {1e-3, 1e-6, 1e-9, 0} below a gap
  and a cutoff inside a degenerate triple give the right energies and subspaces;
* O2: block widths that shrink and grow (8, 3, 1, 12, 2) through the chain, dense
  and fused paths, offset column views whose neighbours stay untouched, a vector
  against a one-column block.  The library's scratch is internal (the reference's
  test fills H.buf1/buf2 with NaN and 1e300 by hand); here every call after the
  first runs on scratch holding the previous, differently shaped call's values.

Every returned pair is checked: energies, projector onto the expected subspace and
fresh residuals against the oracle's chain apply."""
import numpy as np
import pytest

from oracle import kfc_oracle as O
from tests.helpers import oracle_from_kfc_terms, random_ops, to_kfc, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _ortho(rng, n):
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    return Q * np.sign(np.diag(R))


def _dense_axis(M):
    import kfc
    return kfc.KronFreeOperator([M.shape[0]], [kfc.AxisTerm(1.0, [(0, kfc.DenseOp(M))])])


def _with_spectrum(Q, lam):
    M = Q @ np.diag(lam) @ Q.T
    return (M + M.T) / 2


def _check_pairs(H, E, V, tol):
    Ho = oracle_from_kfc_terms(H)
    r = O.true_residuals(Ho, E, V)
    assert np.max(r) <= tol, np.max(r)
    assert np.allclose(V.conj().T @ V, np.eye(V.shape[1]), atol=1e-8)


def _solvers():
    import kfc
    return {"lobpcg": lambda tol, **kw: kfc.LOBPCG(tol=tol, **kw),
            "chebyshev": lambda tol, **kw: kfc.ChebyshevFilter(tol=tol, extra=10, **kw)}


@pytest.mark.parametrize("solver", ["lobpcg", "lockstep"])
def test_s1_exact_and_wrong_subspace_seeds(solver):
    import kfc
    d = np.linspace(0.0, 20.0, 40)
    H = kfc.KronFreeOperator([40], [kfc.AxisTerm(1.0, [(0, kfc.DiagonalOp(d))])])
    Id = np.eye(40)
    s = kfc.LOBPCG(tol=1e-10)
    E0, V0 = kfc.solve(s, H, 3)
    for X0 in (Id[:, :3], Id[:, 5:8], V0):
        if solver == "lobpcg":
            E, V = kfc.solve(s, H, 3, X0=X0)
        else:
            (E, V), = [x for x in kfc.solve_batch(s, [H], 3, X0s=[X0], lockstep=True)[0]]
        assert np.allclose(E, d[:3], atol=1e-8)
        assert np.allclose(V @ V.T, Id[:, :3] @ Id[:, :3].T, atol=1e-8)
        _check_pairs(H, E, V, 1e-9)


@pytest.mark.parametrize("solver", ["lobpcg", "chebyshev", "lanczos"])
@pytest.mark.parametrize("c", [0.0, 2.5])
def test_s3_multiple_of_identity(solver, c):
    """H = c I: valid pairs (every E = c) or an informative rejection naming Dense, as
    the reference's device adjustment allows (test/cuda/test_device.jl:312-324)."""
    import kfc
    N = 60
    H = kfc.KronFreeOperator([N], [], constant=c)
    s = {"lobpcg": kfc.LOBPCG(), "chebyshev": kfc.ChebyshevFilter(extra=10), "lanczos": kfc.Lanczos()}[solver]
    try:
        E, V = kfc.solve(s, H, 3)
    except (kfc.BreakdownError, kfc.ConvergenceError, kfc.ArgumentError) as e:
        assert "Dense" in str(e) or "converge" in str(e), str(e)
    else:
        assert np.allclose(E, c, atol=1e-12)
        _check_pairs(H, E, V, 1e-12)
    Ed, _ = kfc.solve(kfc.Dense(), H, 3)
    assert np.allclose(Ed, c)


@pytest.mark.parametrize("solver", ["lobpcg", "chebyshev", "lockstep"])
def test_s3_tight_clusters(solver):
    import kfc
    rng = np.random.default_rng(43)
    N = 40
    Q = _ortho(rng, N)
    rest = list(np.linspace(1.0, 10.0, N - 3))
    deltas = (1e-3, 1e-6, 1e-9, 0.0)
    Hs = [_dense_axis(_with_spectrum(Q, [0.0, dl, 2 * dl] + rest)) for dl in deltas]
    if solver == "lockstep":
        outs = [x for x in kfc.solve_batch(kfc.LOBPCG(tol=1e-9), Hs, 3, lockstep=True)[0]]
    else:
        outs = [kfc.solve(_solvers()[solver](1e-9), H, 3) for H in Hs]
    P3 = Q[:, :3] @ Q[:, :3].T
    for dl, H, (E, V) in zip(deltas, Hs, outs):
        assert np.allclose(E, [0.0, dl, 2 * dl], atol=1e-8), (dl, E)
        assert np.allclose(V @ V.T, P3, atol=1e-7), dl
        _check_pairs(H, E, V, 1e-8)


@pytest.mark.parametrize("solver", ["lobpcg", "chebyshev", "lockstep"])
def test_s3_cutoff_inside_degenerate_triple(solver):
    import kfc
    rng = np.random.default_rng(44)
    N = 40
    Q = _ortho(rng, N)
    H = _dense_axis(_with_spectrum(Q, [0.0, 1.0, 1.0, 1.0] + list(np.linspace(2.0, 10.0, N - 4))))
    if solver == "lockstep":
        (E, V), = kfc.solve_batch(kfc.LOBPCG(tol=1e-9), [H], 2, lockstep=True)[0]
    else:
        E, V = kfc.solve(_solvers()[solver](1e-9), H, 2)
    assert np.allclose(E, [0.0, 1.0], atol=1e-8)
    assert np.linalg.norm(V @ (V.T @ Q[:, 0]) - Q[:, 0]) < 1e-7
    Pb = Q[:, 1:4] @ Q[:, 1:4].T
    assert np.linalg.norm(Pb @ V[:, 1] - V[:, 1]) < 1e-7
    _check_pairs(H, E, V, 1e-8)


def test_lanczos_resolvable_cluster():
    """A single-vector Krylov method resolves a cluster only as far as its gaps allow
    (KrylovKit's Lanczos, like this one, cannot see an exact multiplicity): delta =
    1e-3 below a gap of 1."""
    import kfc
    rng = np.random.default_rng(45)
    N = 80
    Q = _ortho(rng, N)
    H = _dense_axis(_with_spectrum(Q, [0.0, 1e-3, 2e-3] + list(np.linspace(1.0, 10.0, N - 3))))
    E, V = kfc.solve(kfc.Lanczos(tol=1e-10), H, 3)
    assert np.allclose(E, [0.0, 1e-3, 2e-3], atol=1e-9)
    assert np.allclose(V @ V.T, Q[:, :3] @ Q[:, :3].T, atol=1e-6)
    _check_pairs(H, E, V, 1e-9)


@pytest.mark.parametrize("path", ["chain", "dense_pair", "fused"])
def test_o2_shrinking_growing_widths_and_views(path):
    import kfc
    from kfc import configs
    rng = np.random.default_rng(45)
    if path == "chain":
        dims = [4, 5, 3]
        terms = random_ops(rng, dims, 4, True)
        H = to_kfc(dims, terms, constant=0.3)
        Ho = to_oracle(dims, terms, constant=0.3)
    elif path == "dense_pair":
        dims = [9, 7, 6]
        mk = lambda n: (lambda A: (A + A.T) / 2)(rng.standard_normal((n, n)))
        H = kfc.KronFreeOperator(dims, [kfc.AxisTerm(1.0, [(1, kfc.DiagonalOp(rng.uniform(0, 2, 7)))]),
                                        kfc.AxisTerm(0.7, [(0, kfc.DenseOp(mk(9))), (1, kfc.DenseOp(mk(7)))]),
                                        kfc.AxisTerm(0.4, [(1, kfc.DenseOp(mk(7))), (2, kfc.DenseOp(mk(6)))])],
                                 block_capacity=8)
        Ho = oracle_from_kfc_terms(H)
    else:
        H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(9), N=20)
        assert H.fused
        Ho = oracle_from_kfc_terms(H)
    dt = torch.complex128 if H.dtype == np.complex128 else torch.float64
    N = H.N
    for w in (8, 3, 1, 12, 2):
        Xp = torch.tensor(rng.standard_normal((w + 3, N + 5)), dtype=torch.float64, device="cuda").to(dt)
        if dt == torch.complex128:
            Xp = Xp + 1j * torch.tensor(rng.standard_normal((w + 3, N + 5)), device="cuda")
        Yp = torch.full((w + 3, N + 5), -5.0, dtype=dt, device="cuda")
        saved = Xp.cpu().numpy().copy()
        X, Y = Xp[1:w + 1, :N], Yp[2:w + 2, :N]        # offset views, row stride N + 5
        H.mul(Y, X)
        ref = O.apply_chain(Ho, Xp[1:w + 1, :N].cpu().numpy().T).T
        got = Y.cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref))), w
        assert np.array_equal(Xp.cpu().numpy(), saved)
        Yh = Yp.cpu().numpy()
        assert np.all(Yh[[0, 1, w + 2]] == -5.0) and np.all(Yh[2:w + 2, N:] == -5.0)
    x = torch.tensor(rng.standard_normal(N), dtype=torch.float64, device="cuda").to(dt)
    y1 = (H @ x).cpu().numpy()
    y2 = (H @ x.reshape(1, N)).cpu().numpy()[0]
    assert np.array_equal(y1, y2)
    Z = torch.zeros((4, N), dtype=dt, device="cuda")
    with pytest.raises(kfc.ArgumentError):
        H.mul(Z[1:3], Z[0:2])                             # overlapping views


@pytest.mark.parametrize("path", ["chain", "dense_pair", "fused"])
def test_o2_stale_scratch(path):
    """test/cuda/test_device.jl:163-207: the scratch filled with NaN, then with 1e300,
    between applies of changing widths at both precisions; every result equals the
    oracle's, and a solve that follows a NaN fill gives valid pairs."""
    import kfc
    from kfc import configs
    rng = np.random.default_rng(46)
    if path == "chain":
        dims = [4, 5, 3]
        terms = random_ops(rng, dims, 4, True)
        H = to_kfc(dims, terms, constant=0.3)
        Ho = to_oracle(dims, terms, constant=0.3)
    elif path == "dense_pair":
        dims = [9, 7, 6]
        mk = lambda n: (lambda A: (A + A.T) / 2)(rng.standard_normal((n, n)))
        H = kfc.KronFreeOperator(dims, [kfc.AxisTerm(1.0, [(1, kfc.DiagonalOp(rng.uniform(0, 2, 7)))]),
                                        kfc.AxisTerm(0.7, [(0, kfc.DenseOp(mk(9))), (1, kfc.DenseOp(mk(7)))]),
                                        kfc.AxisTerm(0.4, [(1, kfc.DenseOp(mk(7))), (2, kfc.DenseOp(mk(6)))])],
                                 block_capacity=8)
        Ho = oracle_from_kfc_terms(H)
    else:
        H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(9), N=20)
        Ho = oracle_from_kfc_terms(H)
    cplx = H.dtype == np.complex128
    for w, fill in ((4, None), (3, float("nan")), (5, 1e300), (2, float("nan"))):
        if fill is not None:
            H.fill_scratch(fill)
        X = rng.standard_normal((w, H.N)) + (1j * rng.standard_normal((w, H.N)) if cplx else 0)
        ref = O.apply_chain(Ho, X.T).T
        for prec, tol in (("f64", 1e-12), ("f32", 2e-5)):
            dt = {("f64", False): torch.float64, ("f32", False): torch.float32,
                  ("f64", True): torch.complex128, ("f32", True): torch.complex64}[(prec, cplx)]
            Xt = torch.tensor(X, device="cuda").to(dt)
            Yt = torch.empty_like(Xt)
            H.mul(Yt, Xt, prec)
            got = Yt.to(torch.complex128 if cplx else torch.float64).cpu().numpy()
            assert np.all(np.isfinite(got)), (path, w, fill, prec)
            assert np.max(np.abs(got - ref)) <= tol * max(1.0, np.max(np.abs(ref))), (path, w, fill, prec)
    if path == "dense_pair":   # Hermitian by construction
        H.fill_scratch(float("nan"))
        E, V = kfc.solve(kfc.LOBPCG(tol=1e-9), H, 2)
        _check_pairs(H, E, V, 1e-8)


def test_no_device_allocation_after_warm_up():
    """test/cuda/test_device.jl:18-36: the operator's applies on a block and on a
    column view of it equal the host result, and after warm-up an apply on the view
    allocates no device memory (the scratch blocks are kept at their largest width)."""
    import kfc
    rng = np.random.default_rng(61)
    cd = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    dims = [40, 30, 3]
    H = kfc.KronFreeOperator(dims, [kfc.AxisTerm(1.0, [(0, kfc.DiagonalOp(np.arange(1, 41).astype(complex)))]),
                                    kfc.AxisTerm(0.2, [(0, kfc.DenseOp(cd(40, 40))), (1, kfc.DenseOp(cd(30, 30)))]),
                                    kfc.AxisTerm(0.1, [(1, kfc.ShiftOp(30, 2, 0.5)),
                                                       (2, kfc.BandedOp(3, [-1, 1], [np.ones(3), np.ones(3)]))])],
                             block_capacity=6)
    Ho = oracle_from_kfc_terms(H)
    Xh = cd(6, 3600)
    Xd = torch.tensor(Xh, device="cuda")
    Yd = torch.empty_like(Xd)
    H.mul(Yd, Xd)
    assert np.max(np.abs(Yd.cpu().numpy() - O.apply_chain(Ho, Xh.T).T)) <= 1e-12 * np.max(np.abs(Yd.cpu().numpy()))
    Xv, Yv = Xd[1:5], Yd[1:5]
    H.mul(Yv, Xv)
    H.mul(Yv, Xv)
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(10):
        H.mul(Yv, Xv)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
    assert np.max(np.abs(Yv.cpu().numpy() - O.apply_chain(Ho, Xh[1:5].T).T)) <= 1e-12 * np.max(np.abs(Yv.cpu().numpy()))
    assert np.allclose(H.diagonal(), np.real(np.diag(O.to_matrix(Ho))), atol=1e-12)
