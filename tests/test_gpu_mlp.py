"""The controller network's HIP kernels (include/qpc_mlp.h, FP32) against the
FP64 oracle on the same parameters: forward, loss and every gradient entry;
training on the device reduces the loss and is reproducible bit for bit."""
import numpy as np
import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu


def test_mlp_forward_and_gradient_match_oracle():
    import qpc_gpu
    print("case: mlp forward/gradient n=300 vs FP64 oracle", flush=True)
    net = qpc_gpu.Mlp(seed=11)
    rng = np.random.default_rng(5)
    P = net.params.astype(np.float64)
    P[P.size - 2:] = (0.3, -0.2)  # non-zero output biases
    P[448:512] = 0.05 * rng.standard_normal(64)
    net.params = P.astype(np.float32)
    P = net.params.astype(np.float64)
    X = rng.standard_normal((300, 7)).astype(np.float32)
    T = rng.standard_normal((300, 2)).astype(np.float32)
    Y = net.forward(X)
    Yr = po.mlp_forward(P, X.astype(np.float64))
    assert np.abs(Y - Yr).max() < 2e-6 * max(1.0, np.abs(Yr).max())
    loss, G = net.loss_grad(X, T)
    lr, Gr = po.mlp_loss_grad(P, X.astype(np.float64), T.astype(np.float64))
    print(f"  loss {loss:.7f} / {lr:.7f}, max |dG| {np.abs(G - Gr).max():.2e} of {np.abs(Gr).max():.2e}", flush=True)
    assert abs(loss - lr) < 1e-5 * lr
    assert np.abs(G - Gr).max() < 1e-5 * np.abs(Gr).max()


def test_mlp_training_reduces_loss_reproducibly():
    import qpc_gpu
    print("case: mlp training n=2048, 200 epochs", flush=True)
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, (2048, 7)).astype(np.float32)
    T = np.stack([np.sin(2 * X[:, 0]) + 0.5 * X[:, 1], X[:, 2] * X[:, 3] - 0.3 * X[:, 6]], axis=1).astype(np.float32)
    runs = []
    for _ in range(2):
        net = qpc_gpu.Mlp(seed=2)
        l0 = net.loss_grad(X, T)[0]
        l1 = net.train(X, T, epochs=200, batch=256, lr=3e-3, seed=4)
        runs.append((l0, l1, net.params))
    (l0, l1, p1), (_, l1b, p2) = runs
    print(f"  loss {l0:.4f} -> {l1:.6f}", flush=True)
    assert l1 < l0 / 20
    assert l1 == l1b and np.array_equal(p1, p2)
