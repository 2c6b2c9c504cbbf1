"""The library's C++ preprocessing (fbh_wbart_inputs / fbh_pbart_inputs, R/fwbart.R:8-27 and R/fpbart.R:7-21) on the
CPU: its quantile functions against closed forms and an independent implementation (scipy), lm's residual standard
error against least squares with aliased columns, mean() against the exact rational mean, and the model matrix by
brute force. The end-to-end known answers (sigest, qchisq and qnorm table values, overrides) are in test_api_cpu.py,
which goes through the same entry points."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest
from scipy import special
from scipy import stats as st

from fastbart_amd import api
from fastbart_amd._lib import lib


def test_qnorm_against_closed_values_and_scipy():
    L = lib()
    assert abs(L.fbh_qnorm(0.975) - 1.959963984540054) < 1e-15
    assert L.fbh_qnorm(0.5) == 0.0
    assert math.isinf(L.fbh_qnorm(0.0)) and L.fbh_qnorm(0.0) < 0 and math.isnan(L.fbh_qnorm(1.5))
    ps = np.concatenate([np.logspace(-300, -1, 300), np.linspace(0.01, 0.99, 397), 1 - np.logspace(-15, -1, 100)])
    for p in ps:
        a, b = L.fbh_qnorm(float(p)), float(st.norm.ppf(p))
        assert abs(a - b) <= 2e-14 * max(1.0, abs(b)), p
    # symmetry: qnorm(1 - p) = -qnorm(p) where 1 - p is exact
    for p in (0.125, 0.25, 0.375, 0.0625):
        assert L.fbh_qnorm(1 - p) == -L.fbh_qnorm(p)


def test_qchisq_against_closed_forms_and_scipy():
    L = lib()
    for p in (1e-12, 1e-6, 0.01, 0.1, 0.5, 0.9, 0.999, 1 - 1e-9):
        # df = 2 is the exponential: qchisq(p, 2) = -2 log(1 - p)
        assert abs(L.fbh_qchisq(p, 2.0) - (-2.0 * math.log1p(-p))) <= 1e-13 * (-2.0 * math.log1p(-p))
        # df = 1 is a squared normal: qchisq(p, 1) = qnorm((1 + p) / 2)^2 = 2 erfinv(p)^2
        ref = 2.0 * float(special.erfinv(p)) ** 2
        assert abs(L.fbh_qchisq(p, 1.0) - ref) <= 1e-11 * ref
    for df in (0.5, 1.0, 3.0, 5.0, 10.0, 30.0, 100.0, 1000.0):
        for p in (1e-8, 1e-3, 0.05, 0.1, 0.3, 0.5, 0.7, 0.95, 0.999):
            a, b = L.fbh_qchisq(p, df), float(st.chi2.ppf(p, df))
            assert abs(a - b) <= 1e-12 * b, (p, df, a, b)
    assert abs(L.fbh_qchisq(0.1, 3.0) - 0.5843743741) < 1e-9      # table value
    assert L.fbh_qchisq(0.0, 3.0) == 0.0 and math.isinf(L.fbh_qchisq(1.0, 3.0))


def test_lm_sigma_with_full_rank_and_aliased_columns():
    rng = np.random.default_rng(7)
    n = 200
    X = rng.standard_normal((n, 5))
    y = X @ rng.standard_normal(5) + rng.standard_normal(n)
    A = np.column_stack([np.ones(n), X])
    r = y - A @ np.linalg.lstsq(A, y, rcond=None)[0]
    pre = api.wbart_inputs(X, y, ntree=50)
    yc = y - pre["fmean"]
    rc = yc - A @ np.linalg.lstsq(A, yc, rcond=None)[0]
    assert pre["lm_rank"] == 6
    assert abs(pre["sigest"] - math.sqrt(rc @ rc / (n - 6))) <= 1e-12 * pre["sigest"]
    # a column that is a combination of two others (and one that repeats the intercept after rm.const keeps it):
    # aliased, the residual degrees of freedom are n - rank
    X2 = np.column_stack([X, 2 * X[:, 0] - X[:, 3], X[:, 1]])
    pre2 = api.wbart_inputs(X2, y, ntree=50)
    assert pre2["lm_rank"] == 6
    assert abs(pre2["sigest"] - pre["sigest"]) <= 1e-10 * pre["sigest"]
    # p >= n: sd(yc)
    pre3 = api.wbart_inputs(rng.standard_normal((10, 12)), y[:10], ntree=10)
    assert pre3["lm_rank"] == 0 and abs(pre3["sigest"] - np.std(y[:10], ddof=1)) < 1e-14


def test_mean_is_the_correctly_rounded_mean():
    """R's mean(): a long double sum and a correction pass. On data whose double sum loses digits (a large common
    offset, magnitudes over six decades) the result is still the exact mean rounded to double."""
    rng = np.random.default_rng(8)
    for trial in range(20):
        y = 1e8 + rng.uniform(-1, 1, 100) * 10.0 ** rng.integers(-3, 4, 100)
        exact = sum(Fraction(float(v)) for v in y) / len(y)
        pre = api.wbart_inputs(rng.random((100, 2)), y, ntree=10, lambda_=1.0)
        got = Fraction(pre["fmean"])
        ulp = Fraction(math.ulp(float(exact)))
        assert abs(got - exact) <= ulp / 2, trial


def test_model_matrix_brute_force():
    rng = np.random.default_rng(9)
    n = 300
    x = np.column_stack([rng.random(n), rng.integers(0, 5, n).astype(float), np.full(n, 2.0), rng.random(n) < 0.3,
                         rng.normal(size=n)])
    X, xi, rmc = api.model_matrix(x, numcut=20)
    assert np.array_equal(rmc, [1, 2, 4, 5])
    # few distinct values: midpoints
    u = np.unique(x[:, 1])
    assert np.array_equal(xi[1], (u[1:] + u[:-1]) / 2) and np.array_equal(xi[2], [0.5])
    # many: numcut equally spaced points strictly inside [min, max]
    for v, col in ((0, x[:, 0]), (3, x[:, 4])):
        c = xi[v]
        lo, hi = col.min(), col.max()
        assert len(c) == 20 and lo < c[0] and c[-1] < hi
        d = np.diff(c)
        assert np.allclose(d, (hi - lo) / 21, rtol=1e-12, atol=0)
        assert abs(c[0] - (lo + (hi - lo) / 21)) <= 1e-15 * max(1, abs(lo))
    pre = api.wbart_inputs(x, rng.standard_normal(n), numcut=20)
    for v in range(4):
        brute = np.array([sum(1 for t in xi[v] if t <= xv) for xv in X[:, v]])
        assert np.array_equal(pre["q"][v], brute)
    assert np.array_equal(pre["numcut"], [20, 4, 1, 20])


def test_preprocessing_refusals():
    rng = np.random.default_rng(10)
    x = rng.random((50, 2))
    with pytest.raises(ValueError, match="nondecreasing"):
        api.wbart_inputs(x, rng.random(50), xinfo=[np.array([0.5, 0.2]), np.array([0.5])])
    with pytest.raises(ValueError, match="non-finite"):
        bad = x.copy(); bad[3, 1] = np.nan
        api.wbart_inputs(bad, rng.random(50))
    with pytest.raises(ValueError, match="0/1"):
        api.pbart_inputs(x, np.arange(50) % 3)
    with pytest.raises(ValueError, match="predictors"):
        api.wbart_inputs(np.full((50, 3), 1.0), rng.random(50))      # every column constant
