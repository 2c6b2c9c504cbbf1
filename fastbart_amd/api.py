"""The R entry points fwbart() / fpbart() on the device sampler, returning the components of R's wbart / pbart objects.

Mirrors R/fwbart.R:4-45 and R/fpbart.R:4-41 in three steps: preprocessing (``wbart_inputs`` / ``pbart_inputs``:
constant-column removal, cutpoints, quantization, sigest / lambda / tau, the prior table), the sampler call (the
drop-in ``bart_fit_cpu`` with the internal generator, or one batched launch for many data sets), and the
post-processing (``wbart_object`` / ``pbart_object``: treedraws text, varcount, yhat.train in tree order, + fmean or
+ binaryOffset, pnorm, varprob). BART::bartModelMatrix is not available here: cutpoints default to its equal grid
between each column's min and max (level midpoints for columns with few values); results equal R's for the same
cutpoints, which ``xinfo`` passes in.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import stats as _st

from . import sampler
from . import workloads as W

MAX_N, MAX_P, MAX_NTREE, MAX_NUMCUT = 65536, 32, 200, 255


def model_matrix(x, numcut=100, xinfo=None, rm_const=True):
    """(X with constant columns removed, per-column cutpoints, rm.const = kept columns 1-based)."""
    X = np.asarray(x, dtype=np.float64)
    if X.ndim == 1:
        X = X[:, None]
    keep = [v for v in range(X.shape[1]) if not rm_const or len(np.unique(X[:, v])) > 1]
    if xinfo is not None:
        xi = [np.asarray(xinfo[v], np.float64) for v in keep]
    else:
        xi = [W.make_cutpoints(X[:, v], numcut) for v in keep]
    return X[:, keep], xi, np.array([v + 1 for v in keep], np.int32)


def _check(name, n, p, ntree, xi, extra):
    """The conditions under which the R wrappers leave for BART (R/fwbart.R:11, R/fpbart.R:11) raise here."""
    if extra:
        raise ValueError(f"{name}: unsupported arguments {sorted(extra)}")
    if n > MAX_N or p > MAX_P or p < 1 or ntree > MAX_NTREE or any(len(c) > MAX_NUMCUT for c in xi):
        raise ValueError(f"{name}: unsupported settings (n={n} > {MAX_N}, p={p} outside 1..{MAX_P}, "
                         f"ntree={ntree} > {MAX_NTREE} or numcut > {MAX_NUMCUT}); R falls back to BART here")


def _quantize(X, xi):
    return np.stack([W.quantize(X[:, v], xi[v]) for v in range(X.shape[1])])


def wbart_inputs(x_train, y_train, w=None, sigest=None, sigdf=3.0, sigquant=0.90, k=2.0, power=2.0, base=0.95,
                 sigmaf=None, lambda_=None, fmean=None, ntree=200, numcut=100, xinfo=None, rm_const=True, **extra):
    """Sampler inputs of R/fwbart.R:8-27 (plus the cutpoints, rm.const and fmean the object keeps)."""
    X, xi, rmc = model_matrix(x_train, numcut, xinfo, rm_const)
    y = np.asarray(y_train, np.float64)
    n, p = len(y), X.shape[1]
    _check("fwbart", n, p, ntree, xi, extra)
    fmean = float(np.mean(y)) if fmean is None else float(fmean)
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    yc = y - fmean
    nu = float(sigdf)
    if lambda_ is None:
        if sigest is None:
            sigest = W.ols_sigma(X, yc) if p < n else float(np.std(yc, ddof=1))
        lam = sigest * sigest * float(_st.chi2.ppf(1.0 - sigquant, nu)) / nu
    else:
        lam = float(lambda_)
        sigest = math.sqrt(lam)
    tau = (float(np.max(yc)) - float(np.min(yc))) / (2 * k * math.sqrt(ntree)) if sigmaf is None \
        else float(sigmaf) / math.sqrt(ntree)
    return dict(q=_quantize(X, xi), y=yc.astype(np.float32), wt=(1.0 / w ** 2).astype(np.float32),
                numcut=np.array([len(c) for c in xi], np.int32), pgtab=W.pgtab(base, power), tau=tau, nu=nu,
                sigest=float(sigest), offset=0.0, cutpoints=xi, fmean=fmean, ntree=int(ntree),
                **{"lambda": lam, "rm.const": rmc})


def pbart_inputs(x_train, y_train, k=2.0, power=2.0, base=0.95, binary_offset=None, ntree=50, numcut=100,
                 xinfo=None, rm_const=True, **extra):
    """Sampler inputs of R/fpbart.R:7-20: latent probit, sigma fixed at 1, offset = qnorm(mean(y))."""
    y01 = np.asarray(y_train).astype(np.int32)
    if binary_offset is None:
        binary_offset = float(_st.norm.ppf(np.mean(y01)))
    X, xi, rmc = model_matrix(x_train, numcut, xinfo, rm_const)
    n, p = len(y01), X.shape[1]
    _check("fpbart", n, p, ntree, xi, extra)
    return dict(q=_quantize(X, xi), y=np.zeros(n, np.float32), wt=np.ones(n, np.float32),
                numcut=np.array([len(c) for c in xi], np.int32), pgtab=W.pgtab(base, power),
                tau=3.0 / (k * math.sqrt(ntree)), nu=3.0, sigest=1.0, offset=float(binary_offset), ybin=y01,
                cutpoints=xi, ntree=int(ntree), **{"lambda": 1.0, "rm.const": rmc})


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
