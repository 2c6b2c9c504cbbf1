"""Synthetic inputs for the sampler workloads, generated deterministically with numpy's PCG64.

Each builder returns a *problem* dict in the form the C-ABI takes (after the R layer's preprocessing,
R/fwbart.R:18-32 and R/fpbart.R:7,17-26): ``q`` [p][n] uint8 cutpoint ranks (x < xi[c] <=> q <= c), ``y`` float32
centred outcome, ``wt`` float32 = 1/w^2, ``numcut`` int32 [p], ``pgtab`` float32 [64], scalars ``tau``, ``lambda``,
``nu``, ``sigest``, ``offset`` and for probit ``ybin``. It also keeps the raw ``x`` columns and ``cutpoints`` for
the predictor.

Workloads (data seed = 1000 * workload + chain; sampler seed = 1 for single chains, 1 + chain in a batch):
  fwbart_n4096      n=4,096, p=16, 100 cutpoints, m=200
  fpbart_n24576     probit, n=24,576, p=3, m=50
  batch512          512 chains, n_i in [3091, 4573], p_i in [15, 19], 255 cutpoints, m=200
  maxshape64        64 chains, n=65,536, p=32 (24 continuous + 8 integer columns), m=200, node cap 64
"""
from __future__ import annotations

import math

import numpy as np
from scipy import stats as _st

NODES = 64


def pgtab(base: float = 0.95, power: float = 2.0) -> np.ndarray:
    """Tree-prior split probabilities by depth, float-rounded as the glue passes them (src/glue.cpp:16)."""
    d = np.arange(NODES, dtype=np.float64)
    return (base / (1.0 + d) ** power).astype(np.float32)


def grid_cutpoints(x: np.ndarray, numcut: int) -> np.ndarray:
    """Equally spaced cutpoints strictly inside [min, max]: c_k = min + (k+1)(max-min)/(numcut+1)."""
    lo, hi = float(np.min(x)), float(np.max(x))
    k = np.arange(1, numcut + 1, dtype=np.float64)
    return lo + k * (hi - lo) / (numcut + 1)


def level_cutpoints(x: np.ndarray) -> np.ndarray:
    """Midpoints between consecutive distinct values (integer / binary columns)."""
    u = np.unique(x)
    return (u[1:] + u[:-1]) / 2.0


def make_cutpoints(x: np.ndarray, numcut: int) -> np.ndarray:
    """Columns with at most numcut+1 distinct values get level midpoints, others an equal grid."""
    u = np.unique(x)
    if len(u) <= 1:
        return np.zeros(0)
    if len(u) <= numcut + 1:
        return level_cutpoints(x)
    return grid_cutpoints(x, numcut)


def quantize(x: np.ndarray, xi: np.ndarray) -> np.ndarray:
    """q = number of cutpoints <= x (R's findInterval), so that x < xi[c] <=> q <= c."""
    return np.searchsorted(xi, x, side="right").astype(np.uint8)


def ols_sigma(X: np.ndarray, y: np.ndarray) -> float:
    """Residual standard error of y ~ 1 + X (the lm() estimate R/fwbart.R:24 uses)."""
    n, p = X.shape
    A = np.column_stack([np.ones(n), X])
    coef, _, rank, _ = np.linalg.lstsq(A, y, rcond=None)
    r = y - A @ coef
    return float(np.sqrt(r @ r / (n - rank)))  # lm() drops aliased columns: residual df = n - rank


def wbart_problem(X: np.ndarray, y: np.ndarray, w: np.ndarray | None = None, ntree: int = 200, numcut: int = 100,
                  k: float = 2.0, power: float = 2.0, base: float = 0.95, sigdf: float = 3.0,
                  sigquant: float = 0.90, cutpoints=None) -> dict:
    """Preprocessing of R/fwbart.R:18-32 for a numeric matrix X [n][p]."""
    X = np.asarray(X, np.float64)
    n, p = X.shape
    y = np.asarray(y, np.float64)
    w = np.ones(n) if w is None else np.asarray(w, np.float64)
    xi = cutpoints if cutpoints is not None else [make_cutpoints(X[:, v], numcut) for v in range(p)]
    q = np.stack([quantize(X[:, v], xi[v]) for v in range(p)])
    fmean = float(np.mean(y))
    yc = y - fmean
    sigest = ols_sigma(X, yc) if p < n else float(np.std(yc, ddof=1))
    lam = sigest * sigest * float(_st.chi2.ppf(1.0 - sigquant, sigdf)) / sigdf
    tau = (float(np.max(yc)) - float(np.min(yc))) / (2 * k * math.sqrt(ntree))
    return dict(q=q, y=yc.astype(np.float32), wt=(1.0 / w ** 2).astype(np.float32),
                numcut=np.array([len(c) for c in xi], np.int32), pgtab=pgtab(base, power),
                tau=float(np.float32(tau)), **{"lambda": float(np.float32(lam))}, nu=float(sigdf),
                sigest=float(np.float32(sigest)), offset=0.0, x=X, cutpoints=xi, fmean=fmean, ntree=ntree)


def pbart_problem(X: np.ndarray, y01: np.ndarray, ntree: int = 50, numcut: int = 100, k: float = 2.0,
                  power: float = 2.0, base: float = 0.95, cutpoints=None) -> dict:
    """Preprocessing of R/fpbart.R:7,17-26: latent probit with sigma fixed at 1."""
    X = np.asarray(X, np.float64)
    n, p = X.shape
    y01 = np.asarray(y01).astype(np.int32)
    xi = cutpoints if cutpoints is not None else [make_cutpoints(X[:, v], numcut) for v in range(p)]
    q = np.stack([quantize(X[:, v], xi[v]) for v in range(p)])
    offset = float(_st.norm.ppf(np.mean(y01)))
    tau = 3.0 / (k * math.sqrt(ntree))
    return dict(q=q, y=np.zeros(n, np.float32), wt=np.ones(n, np.float32),
                numcut=np.array([len(c) for c in xi], np.int32), pgtab=pgtab(base, power),
                tau=float(np.float32(tau)), **{"lambda": 1.0}, nu=3.0, sigest=1.0, offset=float(np.float32(offset)),
                ybin=y01, x=X, cutpoints=xi, ntree=ntree)


def friedman_like(x: np.ndarray) -> np.ndarray:
    """10 sin(pi x1 x2) + 20 (x3 - 0.5)^2 + 10 x4 + 5 x5."""
    return (10 * np.sin(np.pi * x[:, 0] * x[:, 1]) + 20 * (x[:, 2] - 0.5) ** 2 + 10 * x[:, 3] + 5 * x[:, 4])


def fwbart_problem(n: int = 4096, p: int = 16, numcut: int = 100, ntree: int = 200, seed: int = 0,
                   sigma: float = 1.0) -> dict:
    """Single weighted chain: x ~ U(0,1)^p, y = friedman_like + N(0, sigma^2), w ~ U(0.5, 2)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.uniform(0.0, 1.0, size=(n, p))
    y = friedman_like(x) + sigma * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, size=n)
    xi = [grid_cutpoints(x[:, v], numcut) for v in range(p)]
    return wbart_problem(x, y, w, ntree=ntree, numcut=numcut, cutpoints=xi)


def workload_fwbart_n4096() -> dict:
    return fwbart_problem(4096, 16, 100, 200, seed=0)


def workload_fpbart_n24576(n: int = 24576) -> dict:
    rng = np.random.Generator(np.random.PCG64(1000))
    x1 = rng.uniform(0.0, 1.0, n)
    x2 = (rng.uniform(0.0, 1.0, n) < 0.4).astype(np.float64)
    x3 = rng.integers(0, 52, n).astype(np.float64)
    eta = -0.8 + 1.2 * x1 - 0.6 * x2 + 0.02 * x3
    y01 = (rng.uniform(0.0, 1.0, n) < _st.norm.cdf(eta)).astype(np.int32)
    X = np.column_stack([x1, x2, x3])
    xi = [grid_cutpoints(x1, 100), level_cutpoints(x2), level_cutpoints(x3)]
    return pbart_problem(X, y01, ntree=50, numcut=100, cutpoints=xi)


def batch512_chain(c: int) -> dict:
    """Chain c of the 512-chain batch: n ~ U{3091..4573}, p ~ U{15..19}, 255 cutpoints, noise sd ~ U(0.5, 2)."""
    rng = np.random.Generator(np.random.PCG64(2000 + c))
    n = int(rng.integers(3091, 4574))
    p = int(rng.integers(15, 20))
    s = float(rng.uniform(0.5, 2.0))
    x = rng.uniform(0.0, 1.0, size=(n, p))
    y = friedman_like(x) + s * rng.standard_normal(n)
    w = rng.uniform(0.5, 2.0, size=n)
    xi = [grid_cutpoints(x[:, v], 255) for v in range(p)]
    return wbart_problem(x, y, w, ntree=200, numcut=255, cutpoints=xi)


def workload_batch512(nchains: int = 512) -> list:
    return [batch512_chain(c) for c in range(nchains)]


def _random_step(rng, x, depth):
    """Random axis-aligned tree of the given depth on the columns of x; returns f(x) with N(0,1) leaves."""
    n, p = x.shape
    out = np.zeros(n)

    def rec(mask, d):
        if d == 0 or mask.sum() < 2:
            out[mask] = rng.standard_normal()
            return
        v = int(rng.integers(0, p))
        col = x[mask, v]
        cut = float(rng.uniform(np.min(col), np.max(col)))
        left = mask & (x[:, v] < cut)
        rec(left, d - 1)
        rec(mask & ~left, d - 1)

    rec(np.ones(n, bool), depth)
    return out


def maxshape64_chain(c: int, n: int = 65536) -> dict:
    """Chain c of the max-shape batch: 24 U(0,1) columns (255 cuts) + 8 integer columns (4..20 levels)."""
    rng = np.random.Generator(np.random.PCG64(3000 + c))
    xc = rng.uniform(0.0, 1.0, size=(n, 24))
    levels = rng.integers(4, 21, size=8)
    xi_ = np.column_stack([rng.integers(0, L, n) for L in levels]).astype(np.float64)
    x = np.column_stack([xc, xi_])
    f = np.zeros(n)
    for _ in range(8):
        f += _random_step(rng, x, int(rng.integers(1, 5)))
    y = f + rng.standard_normal(n)
    w = np.ones(n)
    xi = [grid_cutpoints(xc[:, v], 255) for v in range(24)] + [level_cutpoints(xi_[:, v]) for v in range(8)]
    return wbart_problem(x, y, w, ntree=200, numcut=255, cutpoints=xi)


def small_problem(n: int, p: int, numcut: int, seed: int, weights: str = "uniform", probit: bool = False,
                  ntree: int = 20) -> dict:
    """Small cases for parity tests: few cutpoints, binary columns, weights far from 1, n not a multiple of 16."""
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.uniform(0.0, 1.0, size=(n, p))
    if p >= 2:
        x[:, 1] = (x[:, 1] < 0.5).astype(np.float64)          # a binary column
    xi = [grid_cutpoints(x[:, v], numcut) if len(np.unique(x[:, v])) > numcut + 1 else level_cutpoints(x[:, v])
          for v in range(p)]
    f = friedman_like(np.column_stack([x, np.zeros((n, max(0, 5 - p)))])[:, :max(p, 5)])
    if probit:
        y01 = (rng.uniform(0, 1, n) < _st.norm.cdf((f - f.mean()) / (f.std() + 1e-12))).astype(np.int32)
        return pbart_problem(x, y01, ntree=ntree, cutpoints=xi)
    y = f + rng.standard_normal(n)
    if weights == "uniform":
        w = rng.uniform(0.5, 2.0, size=n)
    elif weights == "extreme":
        w = np.exp(rng.uniform(-2.5, 2.5, size=n))
    else:
        w = np.ones(n)
    return wbart_problem(x, y, w, ntree=ntree, numcut=numcut, cutpoints=xi)
