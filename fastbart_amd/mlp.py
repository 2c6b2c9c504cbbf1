"""Host-side handle of the learned proposal's network (fbh_mlp_* in include/fastbart_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr


class MLP:
    """d_in -> h1 -> h2 -> d_out, ReLU; BF16 weights/activations, FP32 accumulation (MFMA), FP32 Adam masters."""

    def __init__(self, dims, max_batch=512, seed=1, device=0):
        self.dims = tuple(int(d) for d in dims)
        arr = (C.c_int * 4)(*self.dims)
        h = C.c_void_p()
        check(lib().fbh_mlp_create(arr, int(max_batch), int(seed) & 0xFFFFFFFF, device, C.byref(h)), "fbh_mlp_create")
        self.h = h
        self.max_batch = int(max_batch)
        self.nparams = int(lib().fbh_mlp_nparams(h))

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        check(lib().fbh_mlp_params(self.h, ptr(out), None), "fbh_mlp_params")
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(lib().fbh_mlp_params(self.h, None, ptr(p)), "fbh_mlp_params")

    def unpack(self, p):
        """[W1, b1, W2, b2, W3, b3] views of a packed parameter vector."""
        d, out, o = self.dims, [], 0
        for l in range(3):
            w = p[o:o + d[l] * d[l + 1]].reshape(d[l], d[l + 1])
            o += d[l] * d[l + 1]
            b = p[o:o + d[l + 1]]
            o += d[l + 1]
            out += [w, b]
        return out

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        z = np.zeros((x.shape[0], self.dims[3]), np.float32)
        check(lib().fbh_mlp_forward(self.h, ptr(x), x.shape[0], ptr(z)), "fbh_mlp_forward")
        return z

    def grad(self, x, target, mask=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.int32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        loss = C.c_float()
        g = np.zeros(self.nparams, np.float32)
        check(lib().fbh_mlp_grad(self.h, ptr(x), ptr(t), ptr(m), x.shape[0], C.byref(loss), ptr(g)), "fbh_mlp_grad")
        return loss.value, g

    def train(self, x, target, mask=None, steps=100, batch=256, lr=1e-3, seed=1):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(target, dtype=np.int32)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        losses = np.zeros(steps, np.float32)
        check(lib().fbh_mlp_train(self.h, ptr(x), ptr(t), ptr(m), x.shape[0], int(steps), int(batch), float(lr),
                                  int(seed) & 0xFFFFFFFF, ptr(losses)), "fbh_mlp_train")
        return losses

    def timing(self):
        ms, k = C.c_double(), C.c_int()
        check(lib().fbh_mlp_timing(self.h, C.byref(ms), C.byref(k)), "fbh_mlp_timing")
        return ms.value, k.value

    def close(self):
        if getattr(self, "h", None):
            lib().fbh_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _time_forward(batch=512, reps=20):
    """Average device time of the proposal network's forward pass (513 -> 256 -> 256 -> 8160) on a fixed batch."""
    net = MLP((513, 256, 256, 8160), max_batch=batch, seed=1)
    x = np.random.default_rng(0).standard_normal((batch, 513)).astype(np.float32)
    for _ in range(3 + reps):
        net.forward(x)
    ms, k = net.timing()
    flop = 2 * batch * (513 * 256 + 256 * 256 + 256 * 8160)
    print(f"proposal network forward, batch {batch}: {ms:.4f} ms average over {k} calls, "
          f"{flop / ms / 1e9:.1f} TFLOP/s ({flop / 1e9:.3f} GFLOP per call)", flush=True)


if __name__ == "__main__":
    _time_forward()
