"""The BF16-MFMA MLP kernels against the FP64 CPU oracle (oracle.mlp_*)."""
import numpy as np
import pytest

from oracle import kfc_oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _bf16(x):
    """Round to bfloat16 (nearest even), as the kernels round their GEMM operands."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_forward_backward_vs_oracle():
    from kfc.mlp import MLP
    rng = np.random.default_rng(0)
    net = MLP((6, 256, 256, 1024), seed=3)
    params = [p.astype(np.float64) for p in net.get_params()]
    X = rng.uniform(-1, 1, (96, 6))
    Y = net.forward(X)
    Yo, cache = O.mlp_forward(params, X)
    rel = np.linalg.norm(Y - Yo) / np.linalg.norm(Yo)
    assert rel < 2e-2, rel              # BF16 operands, FP32 accumulation
    dY = rng.standard_normal(Y.shape)
    g = net.backward(X, dY)
    go = O.mlp_backward(params, cache, dY)
    for a, b in zip(g, go):
        r = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = np.sum(a * b) / (np.linalg.norm(a) * np.linalg.norm(b))
        assert r < 5e-2 and cos > 0.998, (r, cos)


def test_forward_matches_bf16_rounded_oracle_closely():
    """With the operands rounded to BF16 as the kernel does, only FP32 accumulation remains."""
    from kfc.mlp import MLP
    rng = np.random.default_rng(1)
    net = MLP((6, 64, 32), seed=5)
    W1, b1, W2, b2 = [p.astype(np.float64) for p in net.get_params()]
    X = rng.uniform(-1, 1, (40, 6))
    z1 = _bf16(X) @ _bf16(W1).T + b1
    a1 = O.gelu_tanh(z1)
    y = _bf16(a1) @ _bf16(W2).T + b2
    Y = net.forward(X)
    assert np.max(np.abs(Y - y)) < 2e-3 * np.max(np.abs(y))


def test_training_reduces_loss():
    from kfc.mlp import MLP
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (512, 6))
    T = np.concatenate([np.sin(X @ rng.standard_normal((6, 8))), np.cos(X @ rng.standard_normal((6, 8)))], 1)
    net = MLP((6, 64, 64, 16), seed=1)
    l0 = float(np.mean((net.forward(X) - T) ** 2))
    net.train(X, T, batch=128, steps=400, lr=2e-3)
    l1 = float(np.mean((net.forward(X) - T) ** 2))
    assert l1 < 0.2 * l0, (l0, l1)
