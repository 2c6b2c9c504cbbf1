"""The NumPy reference of the proposal network, pinned by finite differences (CPU)."""
import numpy as np

import mlp_ref


def _params(rng, dims):
    ps = []
    for l in range(3):
        ps.append(rng.standard_normal((dims[l], dims[l + 1])) * np.sqrt(2.0 / dims[l]))
        ps.append(rng.standard_normal(dims[l + 1]) * 0.1)
    return ps


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = (7, 6, 5, 9)
    ps = _params(rng, dims)
    x = rng.standard_normal((4, dims[0]))
    t = np.array([1, 8, 3, 0])
    mask = rng.uniform(size=(4, dims[3])) < 0.7
    mask[np.arange(4), t] = True
    g = mlp_ref.exact_grad(ps, x, t, mask)
    eps = 1e-6
    for l, p in enumerate(ps):
        flat = p.reshape(-1)
        for idx in rng.choice(flat.size, size=min(6, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + eps
            lp = mlp_ref.exact_loss(ps, x, t, mask)
            flat[idx] = old - eps
            lm = mlp_ref.exact_loss(ps, x, t, mask)
            flat[idx] = old
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[l].reshape(-1)[idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5e-3, 3.0e38], np.float32)
    r = mlp_ref.bf16(x)
    assert r[0] == np.float32(1.0) and abs(r[5] - x[5]) <= x[5] * 2 ** -8
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3 * 2^-9 rounds up to 1 + 2^-7; 1 + 2^-9 down to 1
    assert r[1] == np.float32(1.0)
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == np.float32(1.0)
    assert abs(r[4] - x[4]) <= abs(x[4]) * 2 ** -8


def test_bf16_emulating_gradient_is_close_to_exact():
    rng = np.random.default_rng(1)
    dims = (33, 64, 64, 70)
    ps = [p.astype(np.float32) for p in _params(rng, dims)]
    x = rng.standard_normal((16, dims[0])).astype(np.float32)
    t = rng.integers(0, dims[3], 16)
    l1, g1 = mlp_ref.loss_grad(ps, x, t)
    l0 = mlp_ref.exact_loss(ps, x, t)
    g0 = mlp_ref.exact_grad(ps, x, t)
    assert abs(l1 - l0) < 0.02 * abs(l0)
    for a, b in zip(g1, g0):
        assert np.linalg.norm(a - b) < 0.05 * np.linalg.norm(b) + 1e-9
