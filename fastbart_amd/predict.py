"""compile_trees / predict_draw over the C-ABI (R/predict_draw.R:3-18, src/predict.cpp).

``compile_trees(fit)`` parses a fit's treedraws text once into flat breadth-first child-pair arrays on the device;
``predict_draw(ct, newdata, m)`` evaluates draw m (1-based) and equals ``pwbart(newdata, treedraws, mu)[m, ]``:
x >= cut goes right, trees are added in order 1..ntree, then mu. ``predict_draws`` is the batched pwbart path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr


class CompiledTrees:
    """Handle on device-resident flat trees (class ``fastbart_trees`` in R/predict_draw.R:3-11)."""

    def __init__(self, text: str, cutpoints, mu: float = 0.0, rm_const=None, probit: bool = False, device: int = 0):
        flat = np.concatenate([np.asarray(c, np.float64) for c in cutpoints]) if len(cutpoints) else np.zeros(0)
        flat = np.ascontiguousarray(flat if flat.size else np.zeros(1), dtype=np.float64)
        off = np.zeros(len(cutpoints) + 1, np.int32)
        off[1:] = np.cumsum([len(c) for c in cutpoints])
        raw = text.encode()
        h = C.c_void_p()
        check(lib().fbh_compile_trees(raw, len(raw), ptr(flat), ptr(off), len(cutpoints), device, C.byref(h)),
              "fbh_compile_trees")
        self.h = h
        nd, nt, p, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().fbh_trees_info(h, C.byref(nd), C.byref(nt), C.byref(p), C.byref(nn)), "fbh_trees_info")
        self.ndpost, self.ntree, self.p, self.nnodes = nd.value, nt.value, p.value, nn.value
        self.mu = float(mu)
        self.rm_const = rm_const
        self.probit = probit
        self.device = device

    @property
    def rank_form(self) -> bool:
        """True when predictions take the rank-form kernel (sorted, NaN-free cutpoints)."""
        return bool(lib().fbh_trees_rank_form(self.h))

    def timing(self):
        """(average predict-kernel duration in ms, launches) since compilation."""
        ms, k = C.c_double(), C.c_int()
        check(lib().fbh_predict_timing(self.h, C.byref(ms), C.byref(k)), "fbh_predict_timing")
        return ms.value, k.value

    def last_launch(self):
        """(form, rows_per_thread) of the last prediction: form 1 = rank-form kernel, 0 = general, -1 = none yet."""
        f, r = C.c_int(), C.c_int()
        check(lib().fbh_predict_last_launch(self.h, C.byref(f), C.byref(r)), "fbh_predict_last_launch")
        return f.value, r.value

    def arrays(self):
        k, r = self.nnodes, self.ndpost * self.ntree
        var, child, root = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(r, np.int32)
        cut, theta = np.zeros(k), np.zeros(k)
        check(lib().fbh_trees_arrays(self.h, ptr(var), ptr(child), ptr(cut), ptr(theta), ptr(root)),
              "fbh_trees_arrays")
        return dict(var=var, child=child, cut=cut, theta=theta, root=root)

    def _cols(self, newdata):
        x = np.asarray(newdata, dtype=np.float64)
        if x.ndim == 1:
            x = x[:, None]
        if self.rm_const is not None and len(self.rm_const) < x.shape[1]:
            x = x[:, np.asarray(self.rm_const) - 1]
        cols = [np.ascontiguousarray(x[:, v]) for v in range(x.shape[1])]
        arr = (C.c_void_p * max(len(cols), 1))(*[c.ctypes.data for c in cols])
        return cols, arr, x.shape[0]

    def predict_draw(self, newdata, m: int) -> np.ndarray:
        cols, arr, n = self._cols(newdata)
        out = np.zeros(n, np.float64)
        check(lib().fbh_predict_draw(self.h, C.cast(arr, C.c_void_p), n, len(cols), int(m), self.mu, ptr(out)),
              "predict_draw")
        return out

    def predict_draws(self, newdata, draws) -> np.ndarray:
        cols, arr, n = self._cols(newdata)
        d = np.ascontiguousarray(draws, dtype=np.int32)
        out = np.zeros((len(d), n), np.float64)
        check(lib().fbh_predict_draws(self.h, C.cast(arr, C.c_void_p), n, len(cols), ptr(d), len(d), self.mu,
                                      ptr(out)), "predict_draws")
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().fbh_trees_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceRows:
    """Rows uploaded once for repeated prediction (fbh_rows_upload)."""

    def __init__(self, x, device: int = 0):
        x = np.asarray(x, dtype=np.float64)
        self.cols = [np.ascontiguousarray(x[:, v]) for v in range(x.shape[1])]
        arr = (C.c_void_p * max(len(self.cols), 1))(*[c.ctypes.data for c in self.cols])
        h = C.c_void_p()
        check(lib().fbh_rows_upload(C.cast(arr, C.c_void_p), x.shape[0], x.shape[1], device, C.byref(h)),
              "fbh_rows_upload")
        self.h = h
        self.n, self.p = x.shape

    def predict(self, ct: CompiledTrees, draws, out=None):
        d = np.ascontiguousarray(draws, dtype=np.int32)
        if out is None:
            out = np.zeros((len(d), self.n), np.float64)
        check(lib().fbh_predict_rows(ct.h, self.h, ptr(d), len(d), ct.mu, ptr(out)), "fbh_predict_rows")
        return out

    def predict_device(self, ct: CompiledTrees, draws_dev: "DeviceBuffer", ndraws: int, out_dev: "DeviceBuffer"):
        """Enqueue draws_dev[0..ndraws) (device int32, 1-based) into out_dev [ndraws][n] (device float64); no wait."""
        if draws_dev.nbytes < 4 * ndraws or out_dev.nbytes < 8 * ndraws * self.n:
            raise ValueError("predict_device: device buffers too small")
        check(lib().fbh_predict_rows_device(ct.h, self.h, draws_dev.ptr, int(ndraws), ct.mu, out_dev.ptr, None),
              "fbh_predict_rows_device")

    def close(self):
        if getattr(self, "h", None):
            lib().fbh_rows_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory owned by the library (fbh_device_alloc), for the device-side entry points."""

    def __init__(self, nbytes: int, device: int = 0):
        p = C.c_void_p()
        check(lib().fbh_device_alloc(int(nbytes), device, C.byref(p)), "fbh_device_alloc")
        self.ptr, self.nbytes, self.device = p, int(nbytes), device

    @classmethod
    def from_array(cls, a: np.ndarray, device: int = 0) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes, device)
        check(lib().fbh_device_copy(b.ptr, ptr(a), a.nbytes, 0, device), "fbh_device_copy")
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.zeros(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("to_array: larger than the buffer")
        check(lib().fbh_device_copy(ptr(out), self.ptr, out.nbytes, 1, self.device), "fbh_device_copy")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().fbh_device_free(self.ptr, self.device)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_sync(device: int = 0):
    check(lib().fbh_device_sync(device), "fbh_device_sync")


def compile_trees(fit, device: int = 0) -> CompiledTrees:
    """Parse fit['treedraws'] once (R/predict_draw.R:3-11); mu = binaryOffset for probit fits, else fit['mu']."""
    td = fit["treedraws"]
    probit = fit.get("class") == "pbart"
    mu = fit.get("binaryOffset") if probit else fit.get("mu", fit.get("offset", 0.0))
    return CompiledTrees(td["trees"], [np.asarray(c, np.float64) for c in td["cutpoints"]], mu=mu or 0.0,
                         rm_const=fit.get("rm.const"), probit=probit, device=device)


def predict_draw(ct: CompiledTrees, newdata, m: int) -> np.ndarray:
    """f + mu for draw m (1-based) of the compiled trees (R/predict_draw.R:13-18)."""
    return ct.predict_draw(newdata, m)
