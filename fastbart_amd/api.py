"""The R entry points fwbart() / fpbart() on the device sampler, returning the components of R's wbart / pbart objects.

Mirrors R/fwbart.R:4-45 and R/fpbart.R:4-41 in three steps: preprocessing (``wbart_inputs`` / ``pbart_inputs``, which
call the library's C++ preprocessing fbh_wbart_inputs / fbh_pbart_inputs: constant-column removal, cutpoints,
quantization, sigest / lambda / tau, the prior table, the probit offset), the sampler call (the drop-in
``bart_fit_cpu`` with the internal generator, or one batched launch for many data sets), and the post-processing
(``wbart_object`` / ``pbart_object``: treedraws text, varcount, yhat.train in tree order, + fmean or + binaryOffset,
pnorm, varprob). BART::bartModelMatrix is not available here: cutpoints default to its equal grid between each
column's min and max (level midpoints for columns with few values); results equal R's for the same cutpoints, which
``xinfo`` passes in.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
from scipy import stats as _st

from . import sampler
from ._lib import FbhError, Inputs, PrepOpts, lib, ptr

MAX_N, MAX_P, MAX_NTREE, MAX_NUMCUT = 65536, 32, 200, 255
_NA = float("nan")


def _check(name, n, ntree, extra):
    """The conditions under which the R wrappers leave for BART (R/fwbart.R:11, R/fpbart.R:11) raise here; the
    library refuses the rest (more than 32 predictors or 255 cutpoints) with the same message."""
    if extra:
        raise ValueError(f"{name}: unsupported arguments {sorted(extra)}")
    if n > MAX_N or ntree > MAX_NTREE:
        raise ValueError(f"{name}: unsupported settings (n={n} > {MAX_N} or ntree={ntree} > {MAX_NTREE}); R falls back "
                         f"to BART here")


def _columns(x):
    X = np.asarray(x, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    cols = [np.ascontiguousarray(X[:, v]) for v in range(X.shape[1])]
    arr = (C.c_void_p * max(len(cols), 1))(*[c.ctypes.data for c in cols])
    return X, cols, arr


def _xinfo(xinfo, ncol):
    if xinfo is None:
        return None, None
    if len(xinfo) != ncol:
        raise ValueError("xinfo: one cutpoint vector per column of x")
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64) for c in xinfo]) if ncol else np.zeros(0))
    off = np.zeros(ncol + 1, np.int32)
    off[1:] = np.cumsum([len(c) for c in xinfo])
    return (flat if flat.size else np.zeros(1)), off


def _prep(probit, x, y, w=None, xinfo=None, **opts):
    """Run the library's preprocessing; returns (the problem dict the sampler takes, X, the fbh_inputs fields)."""
    X, cols, arr = _columns(x)
    n, ncol = X.shape
    o = PrepOpts()
    lib().fbh_prep_defaults(int(probit), C.byref(o))
    for k, v in opts.items():
        setattr(o, k, _NA if v is None else v)
    xf, xoff = _xinfo(xinfo, ncol)
    h = C.POINTER(Inputs)()
    if probit:
        yv = np.ascontiguousarray(np.asarray(y).astype(np.int32))
        rc = lib().fbh_pbart_inputs(C.cast(arr, C.c_void_p), n, ncol, ptr(yv), ptr(xf), ptr(xoff), C.byref(o),
                                    C.byref(h))
    else:
        yv = np.ascontiguousarray(np.asarray(y, np.float64))
        wv = None if w is None else np.ascontiguousarray(np.asarray(w, np.float64))
        rc = lib().fbh_wbart_inputs(C.cast(arr, C.c_void_p), n, ncol, ptr(yv), ptr(wv), ptr(xf), ptr(xoff),
                                    C.byref(o), C.byref(h))
    if rc:
        msg = lib().fbh_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise FbhError(f"fbh_{'p' if probit else 'w'}bart_inputs failed ({rc}): {msg}")
    try:
        r = h.contents
        p = r.p
        off = np.ctypeslib.as_array(r.cut_off, (p + 1,)).copy()
        cuts = np.ctypeslib.as_array(r.cutpoints, (max(int(off[-1]), 1),)).copy()
        pre = dict(q=np.ctypeslib.as_array(r.q, (p * n,)).reshape(p, n).copy(),
                   y=np.ctypeslib.as_array(r.y, (n,)).copy(), wt=np.ctypeslib.as_array(r.wt, (n,)).copy(),
                   numcut=np.ctypeslib.as_array(r.numcut, (p,)).copy(), pgtab=np.array(r.pgtab[:], np.float32),
                   tau=r.tau_d, nu=r.nu_d, sigest=r.sigest_d, offset=r.offset_d, fmean=r.fmean,
                   cutpoints=[cuts[off[v]:off[v + 1]].copy() for v in range(p)], ntree=int(o.ntree),
                   lm_rank=r.lm_rank,
                   **{"lambda": r.lambda_d, "rm.const": np.ctypeslib.as_array(r.rm_const, (p,)).copy()})
        if probit:
            pre["ybin"] = np.ctypeslib.as_array(r.ybin, (n,)).copy()
    finally:
        lib().fbh_inputs_free(h)
    return pre, X


def model_matrix(x, numcut=100, xinfo=None, rm_const=True):
    """(X with constant columns removed, per-column cutpoints, rm.const = kept columns 1-based), from the library's
    preprocessing (BART::bartModelMatrix's defaults)."""
    X = np.asarray(x, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    pre, _ = _prep(0, X, np.arange(X.shape[0], dtype=np.float64), None, xinfo, numcut=int(numcut),
                   rm_const=int(bool(rm_const)), lambda_=1.0, sigmaf=1.0)
    return X[:, pre["rm.const"] - 1], pre["cutpoints"], pre["rm.const"]


def wbart_inputs(x_train, y_train, w=None, sigest=None, sigdf=3.0, sigquant=0.90, k=2.0, power=2.0, base=0.95,
                 sigmaf=None, lambda_=None, fmean=None, ntree=200, numcut=100, xinfo=None, rm_const=True, **extra):
    """Sampler inputs of R/fwbart.R:8-27 (plus the cutpoints, rm.const and fmean the object keeps), computed by the
    library's fbh_wbart_inputs."""
    _check("fwbart", len(np.asarray(y_train)), ntree, extra)
    pre, _ = _prep(0, x_train, y_train, w, xinfo, ntree=int(ntree), numcut=int(numcut), rm_const=int(bool(rm_const)),
                   k=float(k), power=float(power), base=float(base), sigdf=float(sigdf), sigquant=float(sigquant),
                   sigest=sigest, sigmaf=sigmaf, lambda_=lambda_, fmean=fmean)
    return pre


def pbart_inputs(x_train, y_train, k=2.0, power=2.0, base=0.95, binary_offset=None, ntree=50, numcut=100,
                 xinfo=None, rm_const=True, **extra):
    """Sampler inputs of R/fpbart.R:7-20 (latent probit, sigma fixed at 1, offset = qnorm(mean(y))), computed by the
    library's fbh_pbart_inputs."""
    _check("fpbart", len(np.asarray(y_train)), ntree, extra)
    pre, _ = _prep(1, x_train, y_train, None, xinfo, ntree=int(ntree), numcut=int(numcut),
                   rm_const=int(bool(rm_const)), k=float(k), power=float(power), base=float(base),
                   binary_offset=binary_offset)
    return pre


def _common(res, pre, yhat_train):
    p = pre["q"].shape[0]
    text, vc = sampler.tree_text(res["trees"], p)
    nd = vc.shape[0]
    out = {"treedraws": {"cutpoints": pre["cutpoints"], "trees": text}, "varcount": vc,
           "varprob": np.full((nd, p), 1.0 / p), "moves": np.asarray(res["stats"][:3], np.int32),
           "rm.const": pre["rm.const"]}
    out["varcount.mean"] = vc.mean(axis=0)
    out["varprob.mean"] = out["varprob"].mean(axis=0)
    return out


def wbart_object(res, pre, yhat_train):
    """R/fwbart.R:34-43 on the sampler's buffers; yhat_train = per-draw tree sums (src/glue.cpp:52-56)."""
    out = _common(res, pre, yhat_train)
    fm = pre["fmean"]
    out.update({"sigma": np.asarray(res["sigma"], np.float64), "mu": fm,
                "yhat.train.mean": np.asarray(res["trmean"], np.float64) + fm, "yhat.train": yhat_train + fm,
                "class": "wbart"})
    return out


def pbart_object(res, pre, yhat_train):
    """R/fpbart.R:28-39: no sigma or yhat.train.mean; yhat.train + binaryOffset, prob.train = pnorm of it."""
    out = _common(res, pre, yhat_train)
    off = pre["offset"]
    yh = yhat_train + off
    pt = _st.norm.cdf(yh)
    out.update({"yhat.train": yh, "prob.train": pt, "prob.train.mean": pt.mean(axis=0), "binaryOffset": off,
                "class": "pbart"})
    return out


def fwbart(x_train, y_train, ndpost=1000, nskip=100, seed=1, **kw):
    """wbart object components for a continuous outcome with optional weights ``w`` (R/fwbart.R:4-45).

    Keywords as R's (``lambda`` is ``lambda_``, ``rm.const`` is ``rm_const``, ``binaryOffset`` is ``binary_offset``);
    ``seed`` seeds the sampler's internal generator (R draws it with sample.int, R/fwbart.R:29)."""
    pre = wbart_inputs(x_train, y_train, **kw)
    res = sampler.bart_fit(pre, pre["ntree"], nskip, ndpost, seed)
    return wbart_object(res, pre, sampler.yhat(res["trees"], pre["q"]))


def fpbart(x_train, y_train, ndpost=1000, nskip=100, seed=1, **kw):
    """pbart object components for a binary outcome, probit link (R/fpbart.R:4-41)."""
    pre = pbart_inputs(x_train, y_train, **kw)
    res = sampler.bart_fit(pre, pre["ntree"], nskip, ndpost, seed, probit=1)
    return pbart_object(res, pre, sampler.yhat(res["trees"], pre["q"]))


def fwbart_batch(datasets, ndpost=1000, nskip=100, seeds=None, stats_mode=0, **kw):
    """fwbart on many data sets in one persistent launch (no reference counterpart: the reference runs one chain per
    R process). ``datasets``: list of (x, y) or (x, y, w); shared keywords as fwbart. In the default statistics mode
    each result equals fwbart() on that data set with the same seed."""
    pres = []
    for d in datasets:
        pres.append(wbart_inputs(d[0], d[1], w=d[2] if len(d) > 2 else None, **kw))
    seeds = list(seeds) if seeds is not None else list(range(1, len(pres) + 1))
    ntree = pres[0]["ntree"]
    with sampler.BatchSampler(pres, ntree, nskip, ndpost, seeds, stats_mode=stats_mode) as b:
        b.run(nskip + ndpost)
        return [wbart_object(b.fetch(c), pre, b.yhat(c)) for c, pre in enumerate(pres)]


def fpbart_batch(datasets, ndpost=1000, nskip=100, seeds=None, stats_mode=0, **kw):
    """fpbart on many binary data sets in one persistent launch; ``datasets``: list of (x, y01)."""
    pres = [pbart_inputs(d[0], d[1], **kw) for d in datasets]
    seeds = list(seeds) if seeds is not None else list(range(1, len(pres) + 1))
    with sampler.BatchSampler(pres, pres[0]["ntree"], nskip, ndpost, seeds, probit=1, stats_mode=stats_mode) as b:
        b.run(nskip + ndpost)
        return [pbart_object(b.fetch(c), pre, b.yhat(c)) for c, pre in enumerate(pres)]
