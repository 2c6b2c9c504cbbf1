"""The proposal network on the matrix cores against the NumPy reference: forward, gradients (the reference's own
gradient is pinned by finite differences in test_mlp_reference.py), and training."""
import numpy as np
import pytest

import mlp_ref
from fastbart_amd.mlp import MLP

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,M", [((37, 64, 128, 100), 20), ((513, 256, 256, 8160), 130)])
def test_gpu_forward_matches_reference(dims, M):
    net = MLP(dims, max_batch=256, seed=3)
    p = net.params()
    ps = net.unpack(p)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((M, dims[0])).astype(np.float32)
    z = net.forward(x)
    zr, _ = mlp_ref.forward(ps, x)
    err = np.abs(z - zr)
    print("max abs err", err.max(), "scale", np.abs(zr).max())
    assert err.max() < 2e-2 * (np.abs(zr).max() + 1.0)
    assert np.median(err) < 1e-3 * (np.abs(zr).max() + 1.0)


@pytest.mark.parametrize("dims,M", [((37, 64, 128, 100), 20), ((513, 256, 256, 8160), 96)])
def test_gpu_gradients_match_reference(dims, M):
    net = MLP(dims, max_batch=256, seed=4)
    ps = net.unpack(net.params())
    rng = np.random.default_rng(6)
    x = rng.standard_normal((M, dims[0])).astype(np.float32)
    t = rng.integers(0, dims[3], M)
    mask = rng.uniform(size=(M, dims[3])) < 0.6
    mask[np.arange(M), t] = True
    loss, g = net.grad(x, t, mask)
    lr, gr = mlp_ref.loss_grad(ps, x, t, mask)
    assert abs(loss - lr) < 1e-3 * abs(lr) + 1e-4
    for k, (a, b) in enumerate(zip(net.unpack(g), gr)):
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-12)
        print("block", k, "relative error", rel)
        assert rel < 2e-2, k


def test_gpu_training_learns_a_masked_rule():
    """Targets are a fixed function of the input restricted to each row's allowed classes; Adam on the device drives the
    loss well below the uniform-guess level log(#allowed)."""
    dims = (64, 128, 128, 96)
    rng = np.random.default_rng(7)
    M = 2048
    x = rng.standard_normal((M, dims[0])).astype(np.float32)
    proj = rng.standard_normal((dims[0], dims[3]))
    mask = rng.uniform(size=(M, dims[3])) < 0.5
    score = np.where(mask, x @ proj, -np.inf)
    t = score.argmax(axis=1)
    net = MLP(dims, max_batch=256, seed=8)
    losses = net.train(x, t, mask, steps=400, batch=256, lr=3e-3, seed=9)
    base = float(np.mean(np.log(mask.sum(axis=1))))
    print("uniform", base, "first", losses[:5].mean(), "last", losses[-20:].mean())
    assert losses[-20:].mean() < 0.5 * base
    z = net.forward(x[:256])
    pred = np.where(mask[:256], z, -np.inf).argmax(axis=1)
    assert np.mean(pred == t[:256]) > 0.6
