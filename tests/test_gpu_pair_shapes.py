"""Every launch shape of the dense kept-basis pair kernels against the oracle.

The apply of two-factor dense terms (a branch node's kept-basis ports, the
reference's dense-factor path src/apply.jl:301-340) has several kernel families
and grids, chosen per operator structure by timing on first use.  Each shape is
pinned here with KFC_PAIR_SHAPE and checked on its own: the single-operator apply
against the oracle's per-term chain (FP64 and FP32), and the batched apply of the
lockstep sweep against the one-point solver.  The operators cover unit and
strided p axes (the strided ones also through the LDS-staged kernel), extents that are not multiples of 4 or 16, one and two terms per
group (shared factors merged on the host), every tile size up to 48, and the split form
of one-term groups with a strided p axis (two mode products through scratch)."""
import numpy as np
import pytest

from oracle import kfc_oracle as O
from tests.helpers import oracle_from_kfc_terms, to_kfc, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N_SHAPES = 15          # kPairShapes in dense_pair.hip
TOL64 = 1e-13
TOL32 = 4e-5


def _run(H, X, precision="f64"):
    dt = torch.float64 if precision == "f64" else torch.float32
    # one column goes in as an N-vector, several as a (b, N) block
    Xd = torch.tensor(np.ascontiguousarray(X.T if X.shape[1] > 1 else X[:, 0]), dtype=dt, device="cuda:0")
    Yd = torch.empty_like(Xd)
    H.mul(Yd, Xd, precision)
    H.sync()
    Y = Yd.cpu().numpy()
    return Y.T if Y.ndim == 2 else Y[:, None]


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _pair_operator(dims, pairs, seed):
    rng = np.random.default_rng(seed)
    terms = []
    for t, (p, q) in enumerate(pairs):
        A = rng.standard_normal((dims[p], dims[p]))
        B = rng.standard_normal((dims[q], dims[q]))
        terms.append((float(rng.standard_normal()), [(p, "dense", A), (q, "dense", B)]))
    # a term sharing the first term's A: merged into it on the host (two terms left)
    p, q = pairs[0]
    terms.append((0.7, [(p, "dense", terms[0][1][0][2]), (q, "dense", rng.standard_normal((dims[q], dims[q])))]))
    terms.append((0.3, [(0, "diag", rng.standard_normal(dims[0]))]))
    return to_kfc(dims, terms, constant=0.25), to_oracle(dims, terms, constant=0.25)


CASES = [([40, 40, 30], [(0, 1), (1, 2)]),      # the zero-pi parent's structure (strided p in (1, 2))
         ([3, 37, 29], [(1, 2)]),               # strided p axis, ragged extents
         ([45, 6, 33], [(0, 2), (1, 2)]),        # 48-wide tiles
         ([13, 10, 7, 2], [(0, 1), (2, 3)]),     # 16-wide tiles, four axes
         ([26, 31], [(0, 1)]),                  # 32-wide tiles, one slab per column
         ([18, 11, 26], [(1, 2)]),              # split form (shapes 11-14) as the group writing D .* X
         ([16, 5, 3, 7], [(1, 3)]),             # split form with an axis between p and q
         ([20, 50, 9], [(0, 1), (1, 2)])]       # split form, p extent above 48 (64-wide tiles)


@pytest.mark.parametrize("shape", range(N_SHAPES))
@pytest.mark.parametrize("case", range(len(CASES)))
def test_pair_shape_apply(monkeypatch, shape, case):
    dims, pairs = CASES[case]
    monkeypatch.setenv("KFC_PAIR_SHAPE", str(shape))
    H, Ho = _pair_operator(dims, pairs, seed=100 + case)
    assert not H.fused
    rng = np.random.default_rng(case)
    for ncols in (1, 5):
        X = rng.standard_normal((H.N, ncols))
        Yo = O.apply_chain(Ho, X)
        assert _rel(_run(H, X), Yo) < TOL64, (shape, dims, ncols)
        assert _rel(_run(H, X, "f32"), Yo) < TOL32, (shape, dims, ncols)


@pytest.mark.parametrize("shape", [0, 7, 8, 9, 10, 11, 12, 13, 14])
def test_pair_shape_lockstep(monkeypatch, shape):
    """The batched (one launch for every point) form of the shape in a lockstep sweep:
    each point's pairs match the one-point solver and meet the tolerance."""
    from kfc import configs
    from kfc.sweeps import ZeroPiSweep
    import kfc
    monkeypatch.setenv("KFC_PAIR_SHAPE", str(shape))
    sw = ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=12, m=6, contexts=1)
    pts = [configs.zero_pi_theta(9300 + i) for i in range(4)]
    seq, _ = sw.run_plain(pts[:2])
    lock, _ = sw.run_batched(pts, mode="f64", lockstep=True)
    for (Es, _, _, _), (El, _, sl, _) in zip(seq, lock):
        assert sl["max_residual"] <= 1e-8
        assert np.max(np.abs(Es - El) / np.abs(Es)) < 1e-9
    for th, (El, Vl, _, _) in zip(pts, lock):
        H, _ = sw.parent(th)
        assert np.max(kfc.residuals(H, El, Vl)) <= 1e-8
        assert np.max(O.true_residuals(oracle_from_kfc_terms(H), El, Vl)) <= 1e-8


@pytest.mark.parametrize("shape", [11, 12, 13, 14])
def test_split_form_lockstep(monkeypatch, shape):
    """The split form in the batched launches of a lockstep sweep (one launch per
    stage for every active point, scratch pieces per point from the context): a
    zero-pi parent [16, 16, 10], whose (phi, zeta) group has 16 entries before its p
    axis, so the pinned shape takes the split form there.  Each point's pairs match
    the one-point solver's and have fresh residuals within the tolerance under the
    oracle's chain apply; the step graphs replay bit for bit what the separate
    launches compute."""
    from kfc import configs
    from kfc.sweeps import ZeroPiSweep
    import kfc
    monkeypatch.setenv("KFC_PAIR_SHAPE", str(shape))
    sw = ZeroPiSweep(ncut=12, Nphi=48, Nzeta=10, keep=16, m=6, contexts=1)
    pts = [configs.zero_pi_theta(9400 + i) for i in range(5)]
    H0, _ = sw.parent(pts[0])
    assert list(H0.dims) == [16, 16, 10]
    seq, _ = sw.run_plain(pts[:2])
    monkeypatch.setenv("KFC_GRAPHS", "0")
    eager, _ = sw.run_batched(pts, mode="f64", lockstep=True)
    monkeypatch.setenv("KFC_GRAPHS", "1")
    lock, _ = sw.run_batched(pts, mode="f64", lockstep=True)
    for (Ee, Ve, _, _), (El, Vl, _, _) in zip(eager, lock):
        assert np.array_equal(Ee, El) and np.array_equal(Ve, Vl)
    for (Es, _, _, _), (El, _, sl, _) in zip(seq, lock):
        assert sl["max_residual"] <= 1e-8
        assert np.max(np.abs(Es - El) / np.abs(Es)) < 1e-9
    for th, (El, Vl, _, _) in zip(pts, lock):
        H, _ = sw.parent(th)
        assert np.max(kfc.residuals(H, El, Vl)) <= 1e-8
        assert np.max(O.true_residuals(oracle_from_kfc_terms(H), El, Vl)) <= 1e-8
