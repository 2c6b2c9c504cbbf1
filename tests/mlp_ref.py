"""NumPy reference of the proposal network (test infrastructure): the same arithmetic as the device kernels -- inputs,
weights and hidden activations rounded to bf16, products accumulated in float64 here (float32 on the device) -- and
its exact gradient."""
import numpy as np


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def forward(params, x):
    W1, b1, W2, b2, W3, b3 = params
    xb = bf16(x).astype(np.float64)
    a1 = xb @ bf16(W1).astype(np.float64) + b1
    h1 = bf16(np.maximum(a1, 0).astype(np.float32)).astype(np.float64)
    a2 = h1 @ bf16(W2).astype(np.float64) + b2
    h2 = bf16(np.maximum(a2, 0).astype(np.float32)).astype(np.float64)
    z = h2 @ bf16(W3).astype(np.float64) + b3
    return z, (xb, h1, h2)


def loss_grad(params, x, target, mask=None):
    """Mean masked-softmax cross-entropy and its gradient with respect to (W1, b1, W2, b2, W3, b3). The backward pass
    rounds dZ and the hidden-layer gradients to bf16 where the device does (before the GEMMs that consume them)."""
    W1, b1, W2, b2, W3, b3 = params
    z, (xb, h1, h2) = forward(params, x)
    M = x.shape[0]
    if mask is None:
        mask = np.ones_like(z, dtype=bool)
    mask = mask.astype(bool)
    zm = np.where(mask, z, -np.inf)
    mx = zm.max(axis=1, keepdims=True)
    e = np.where(mask, np.exp(z - mx), 0.0)
    lse = mx[:, 0] + np.log(e.sum(axis=1))
    loss = float(np.mean(lse - z[np.arange(M), target]))
    p = np.where(mask, np.exp(z - lse[:, None]), 0.0)
    dz = p.copy()
    dz[np.arange(M), target] -= 1.0
    dz /= M
    dzb = bf16(dz.astype(np.float32)).astype(np.float64)
    gW3 = h2.T @ dzb
    gb3 = dz.sum(axis=0)
    dh2 = (dzb @ bf16(W3).astype(np.float64).T) * (h2 > 0)
    dh2b = bf16(dh2.astype(np.float32)).astype(np.float64)
    gW2 = h1.T @ dh2b
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2b @ bf16(W2).astype(np.float64).T) * (h1 > 0)
    dh1b = bf16(dh1.astype(np.float32)).astype(np.float64)
    gW1 = xb.T @ dh1b
    gb1 = dh1.sum(axis=0)
    return loss, [gW1, gb1, gW2, gb2, gW3, gb3]


def exact_loss(params, x, target, mask=None):
    """The same loss in float64 without any bf16 rounding (for finite differences)."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    z = h2 @ W3 + b3
    if mask is None:
        mask = np.ones_like(z, dtype=bool)
    zm = np.where(mask, z, -np.inf)
    mx = zm.max(axis=1, keepdims=True)
    lse = mx[:, 0] + np.log(np.where(mask, np.exp(z - mx), 0).sum(axis=1))
    return float(np.mean(lse - z[np.arange(x.shape[0]), target]))


def exact_grad(params, x, target, mask=None):
    """Analytic gradient of exact_loss (float64)."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(p, np.float64) for p in params]
    a1 = x @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    z = h2 @ W3 + b3
    M = x.shape[0]
    if mask is None:
        mask = np.ones_like(z, dtype=bool)
    zm = np.where(mask, z, -np.inf)
    mx = zm.max(axis=1, keepdims=True)
    p = np.where(mask, np.exp(z - mx), 0)
    p /= p.sum(axis=1, keepdims=True)
    dz = p
    dz[np.arange(M), target] -= 1
    dz /= M
    dh2 = (dz @ W3.T) * (a2 > 0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    return [x.T @ dh1, dh1.sum(0), h1.T @ dh2, dh2.sum(0), h2.T @ dz, dz.sum(0)]
