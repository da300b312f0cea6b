"""The boundary around mul! and the solver keywords on the device:

* torch tensors produced by torch kernels go straight into ``H @ x`` and come
  back usable by torch with no explicit synchronization (stream ordering);
* blocks of the wrong element type, device or precision are refused before any
  launch (the reference dispatches on the element type, src/apply.jl:409-416);
* a real operator on complex128 blocks (``_mul_real_operator!``,
  src/apply.jl:397-407) through kfc_apply (KFC_PREC_Z64), against the oracle;
* the ``callback`` keyword of LOBPCG and ChebyshevFilter (src/lobpcg.jl:247-256,
  src/chebyshev.jl:362-363; test/test_solvers.jl:412-426, 775-790): called every
  iteration, locking recorded, exceptions propagate unchanged with no retry."""
import numpy as np
import pytest

from oracle import kfc_oracle as O
from tests.helpers import oracle_from_kfc_terms, random_ops, to_kfc, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def test_mul_ordered_after_torch_kernels():
    """A large tensor written by torch kernels just before the call, read by the
    library at once, and the result read by torch at once (no synchronize)."""
    import kfc
    from kfc import configs
    th = configs.three_fluxonia_theta(3)
    n = 160
    H = configs.three_fluxonia_operator(th, N=n)
    Ho = oracle_from_kfc_terms(H)
    for trial in range(3):
        base = torch.randn(2, H.N, dtype=torch.float64, device="cuda")
        X = (base * 3.0 + 1.0).exp().log() - 1.0          # several torch kernels on the current stream
        Y = H @ X
        Z = (Y * 2.0).cpu().numpy() / 2.0                 # torch reads Y with no explicit sync
        Xh = X.cpu().numpy()
        assert _rel(Z, O.apply_chain(Ho, Xh.T).T) < 1e-13
    # many short applies of freshly allocated tensors in a row
    outs = []
    for k in range(8):
        x = torch.full((H.N,), float(k + 1), dtype=torch.float64, device="cuda")
        outs.append(H @ x)
    ref = O.apply_chain(Ho, np.ones(H.N))
    for k, y in enumerate(outs):
        assert _rel(y.cpu().numpy(), (k + 1) * ref) < 1e-13
    del kfc


def test_mul_refuses_wrong_types():
    import kfc
    rng = np.random.default_rng(5)
    dims = [6, 5]
    terms_r = random_ops(rng, dims, 4, False)
    Hr = to_kfc(dims, terms_r)
    Hc = to_kfc(dims, random_ops(rng, dims, 4, True))
    N = Hr.N
    x64 = torch.randn(2, N, dtype=torch.float64, device="cuda")
    x32 = x64.float()
    z128 = torch.randn(2, N, dtype=torch.complex128, device="cuda")
    with pytest.raises(kfc.ArgumentError):
        Hr.mul(torch.empty_like(x32), x32, "f64")          # FP32 block at FP64 precision
    with pytest.raises(kfc.ArgumentError):
        Hr.mul(torch.empty_like(x64), x64, "bf16")         # unknown precision
    with pytest.raises(kfc.ArgumentError):
        Hc.mul(torch.empty_like(x64), x64, "f64")          # real block on a complex operator
    with pytest.raises(kfc.ArgumentError):
        Hr.mul(torch.empty_like(z128).to(torch.complex64), z128.to(torch.complex64), "f32")
    with pytest.raises(kfc.ArgumentError):
        Hr.mul(torch.empty(2, N, dtype=torch.float64), torch.randn(2, N, dtype=torch.float64), "f64")   # host
    with pytest.raises(kfc.DimensionMismatch):
        Hr.mul(torch.empty(2, N + 1, dtype=torch.float64, device="cuda"),
               torch.randn(2, N + 1, dtype=torch.float64, device="cuda"))
    # the accepted pairs still apply
    y = Hr @ x64
    assert _rel(y.cpu().numpy(), O.apply_chain(to_oracle(dims, terms_r), x64.cpu().numpy().T).T) < 1e-12


@pytest.mark.parametrize("kind", ["fused", "chain", "dense_pair"])
def test_real_operator_on_complex_blocks(kind):
    """_mul_real_operator!: H applied to real and imaginary parts (src/apply.jl:397-407)."""
    import kfc
    from kfc import configs
    rng = np.random.default_rng(11)
    if kind == "fused":
        H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(4), N=24)
        assert H.fused
    elif kind == "chain":
        dims = [7, 6, 5]
        H = to_kfc(dims, random_ops(rng, dims, 5, False))
    else:
        d = [12, 10, 8]
        mk = lambda n: (lambda A: (A + A.T) / 2)(rng.standard_normal((n, n)))
        H = kfc.KronFreeOperator(d, [kfc.AxisTerm(1.0, [(0, kfc.DiagonalOp(rng.uniform(0, 3, 12)))]),
                                     kfc.AxisTerm(0.3, [(0, kfc.DenseOp(mk(12))), (1, kfc.DenseOp(mk(10)))]),
                                     kfc.AxisTerm(-0.2, [(1, kfc.DenseOp(mk(10))), (2, kfc.DenseOp(mk(8)))])])
    assert H.dtype == np.float64
    Ho = oracle_from_kfc_terms(H)
    for b in (1, 3):
        Xh = rng.standard_normal((b, H.N)) + 1j * rng.standard_normal((b, H.N))
        X = torch.tensor(Xh, dtype=torch.complex128, device="cuda")
        if b == 1:
            X = X[0]
        Y = H @ X
        assert Y.dtype == torch.complex128
        ref = O.apply_chain(Ho, Xh.T).T
        got = Y.cpu().numpy().reshape(b, H.N)
        assert _rel(got, ref) < 1e-13
    # offset column views with a row stride above N
    Xp = torch.tensor(rng.standard_normal((4, H.N + 7)) + 1j * rng.standard_normal((4, H.N + 7)),
                      dtype=torch.complex128, device="cuda")
    Yp = torch.full_like(Xp, -5.0)
    Xv, Yv = Xp[1:3, :H.N], Yp[1:3, :H.N]
    H.mul(Yv, Xv)
    ref = O.apply_chain(Ho, Xp[1:3, :H.N].cpu().numpy().T).T
    assert _rel(Yv.cpu().numpy(), ref) < 1e-13
    assert torch.all(Yp[0] == -5.0) and torch.all(Yp[3] == -5.0) and torch.all(Yp[1:3, H.N:] == -5.0)


def test_complex_blocks_on_real_operator_refused_by_host_paths():
    """A complex block with a nonzero imaginary part is never silently truncated."""
    import kfc
    from kfc import configs
    H = configs.three_fluxonia_operator(configs.three_fluxonia_theta(6), N=10)
    E, V = kfc.solve(kfc.Lanczos(tol=1e-9), H, 3)
    with pytest.raises(kfc.ArgumentError):
        kfc.residuals(H, E, V + 1e-3j)
    r = kfc.residuals(H, E, V + 0j)             # exactly real: accepted
    assert np.max(r) < 1e-8
    with pytest.raises(kfc.ArgumentError):
        kfc.solve_batch(kfc.LOBPCG(), [H], 3, X0s=[np.ones((H.N, 3)) * 1j])
    with pytest.raises(kfc.DimensionMismatch):
        kfc.complement_solve(H, torch.zeros(H.N - 1, dtype=torch.float64, device="cuda"),
                             torch.tensor(V.T.copy(), device="cuda"), torch.zeros(1, H.N, dtype=torch.float64,
                                                                                   device="cuda"), [E[0]], 1e-9, 100)


def _oscillators(n):
    """three coupled oscillators on Fock bases, a small real operator with a gapped
    spectrum (in the spirit of test/helpers.jl three_oscillators)."""
    import kfc
    a = np.diag(np.sqrt(np.arange(1, n)), 1)
    x = (a + a.T) / np.sqrt(2)
    terms = [kfc.AxisTerm(w, [(k, kfc.DiagonalOp(np.arange(n) + 0.5))]) for k, w in enumerate((1.0, 1.3, 1.7))]
    terms += [kfc.AxisTerm(0.05, [(0, kfc.DenseOp(x)), (1, kfc.DenseOp(x))]),
              kfc.AxisTerm(0.04, [(1, kfc.DenseOp(x)), (2, kfc.DenseOp(x))])]
    return kfc.KronFreeOperator([n] * 3, terms)


def test_lobpcg_callback_every_iteration():
    import kfc
    H = _oscillators(6)
    w = np.linalg.eigvalsh(O.to_matrix(oracle_from_kfc_terms(H)))
    for mode in ("f64", "mixed"):
        seen = []
        E, V, st = kfc.solve_stats(kfc.LOBPCG(tol=1e-9, callback=lambda info: seen.append(info), mode=mode), H, 6)
        assert np.max(np.abs(E - w[:6])) < 1e-8
        assert len(seen) >= 2
        its = [i.n_iter for i in seen]
        assert its == sorted(its) and len(set(its)) == len(its)
        locked = [i.n_locked for i in seen]
        assert locked == sorted(locked) and locked[-1] == 6 + kfc.solvers.guard_count(kfc.LOBPCG(), H.N, 6)
        assert seen[-1].n_matvec == st["n_matvec"]
        for i in seen:
            # the columns active during the iteration: every one not locked before it
            assert len(i.lambdas) - i.n_locked <= len(i.λs) <= len(i.lambdas)
            assert len(i.λs) == len(i.residual_norms) and len(i.lambdas) == len(i.residuals)
        if mode == "mixed":
            assert seen[0].precision == "f32" and seen[-1].precision == "f64"
    # callbacks of any return type run; the solve is unchanged
    for cb in (lambda info: None, lambda info: info, lambda info: [info.n_iter]):
        E2, _ = kfc.solve(kfc.LOBPCG(tol=1e-9, callback=cb), H, 6)
        assert np.max(np.abs(E2 - w[:6])) < 1e-8


def test_lobpcg_callback_exception_propagates():
    import kfc
    H = _oscillators(6)

    class UserAbort(Exception):
        pass
    calls = []

    def cb(info):
        calls.append(info.n_iter)
        raise UserAbort("user abort")
    for X0 in (None, np.random.default_rng(1).standard_normal((H.N, 6))):
        calls.clear()
        with pytest.raises(UserAbort, match="user abort"):
            kfc.solve(kfc.LOBPCG(callback=cb), H, 6, X0=X0)
        assert calls == [1]        # stopped at the first iteration, no retry from a random block
    # the context is clean afterwards: a solve without callback runs
    E, _ = kfc.solve(kfc.LOBPCG(tol=1e-9), H, 6)
    assert len(E) == 6
    with pytest.raises(kfc.UnsupportedError):
        kfc.solve_batch(kfc.LOBPCG(callback=cb), [H], 6)


def test_chebyshev_callback_locks_at_half_tol():
    """test/test_solvers.jl:412-426: the residuals of locked columns, recorded by the
    callback, are at most tol / 2; iterations counted from 0; the degree is recorded."""
    import kfc
    H = _oscillators(6)
    w = np.linalg.eigvalsh(O.to_matrix(oracle_from_kfc_terms(H)))
    for seed in ("product", "random"):
        locked, its, degs = [], [], []

        def cb(info):
            locked.extend(info.residuals[:info.n_locked])
            its.append(info.iteration)
            degs.append(info.degree)
        s = kfc.ChebyshevFilter(tol=1e-9, degree=50, extra=12, seed=seed, callback=cb)
        E, V = kfc.solve(s, H, 8)
        assert np.max(np.abs(E - w[:8])) < 1e-7
        assert locked and max(locked) <= 0.5 * s.tol
        assert its[0] == 0 and its == list(range(len(its)))
        assert degs[0] == 0 and all(1 <= d <= 50 for d in degs[1:])
    # an exception stops the solve
    with pytest.raises(RuntimeError, match="stop here"):
        kfc.solve(kfc.ChebyshevFilter(tol=1e-9, extra=12, callback=lambda info: (_ for _ in ()).throw(
            RuntimeError("stop here"))), H, 8)
