"""The R-level wrappers' preprocessing and post-processing (R/fwbart.R, R/fpbart.R), checked on the CPU.

Known answers fix each formula: a residual built orthogonal to [1, X] gives sigest exactly, qchisq(0.1, 3) and
qnorm(0.975) are table values, findInterval's convention is checked on values that sit on a cutpoint. The sampler
itself is not called (no GPU here); the object builders are fed the oracle's chain outputs.
"""
import math

import numpy as np
import pytest
from scipy import stats as st

from fastbart_amd import api, sampler


def _orthogonal_residual(X, rng, norm2):
    A = np.column_stack([np.ones(len(X)), X])
    r = rng.standard_normal(len(X))
    r -= A @ np.linalg.lstsq(A, r, rcond=None)[0]
    return r * math.sqrt(norm2 / (r @ r))


def test_model_matrix_drops_constant_columns_and_keeps_one_based_indices():
    rng = np.random.default_rng(0)
    x = np.column_stack([rng.random(50), np.full(50, 3.0), rng.random(50), np.zeros(50)])
    X, xi, rmc = api.model_matrix(x, numcut=10)
    assert X.shape == (50, 2) and np.array_equal(rmc, [1, 3])
    assert np.array_equal(X, x[:, [0, 2]]) and len(xi) == 2
    X2, _, rmc2 = api.model_matrix(x, numcut=10, rm_const=False)
    assert X2.shape == (50, 4) and np.array_equal(rmc2, [1, 2, 3, 4])


def test_quantization_is_find_interval():
    """q = #cutpoints <= x, so x >= xi[c] (goes right at cut c) exactly when q > c (R/fwbart.R:19)."""
    xi = [np.array([0.25, 0.5, 0.75])]
    x = np.array([[0.0], [0.25], [0.49], [0.5], [0.75], [1.0]])
    pre = api.wbart_inputs(x, np.arange(6.0), xinfo=xi, ntree=1)
    assert np.array_equal(pre["q"][0], [0, 1, 1, 2, 3, 3])


def test_sigest_lambda_tau_known_answers():
    rng = np.random.default_rng(1)
    n, p = 60, 3
    X = rng.random((n, p))
    y = 2.0 + X @ np.array([1.0, -2.0, 0.5]) + _orthogonal_residual(X, rng, n - p - 1)   # residual SE exactly 1
    pre = api.wbart_inputs(X, y, ntree=50, k=2.0)
    assert abs(pre["sigest"] - 1.0) < 1e-12
    assert abs(pre["lambda"] - 0.5843744 / 3.0) < 1e-7             # qchisq(0.1, 3) = 0.5843744
    yc = y - y.mean()
    assert abs(pre["tau"] - (yc.max() - yc.min()) / (4 * math.sqrt(50))) < 1e-15
    assert abs(pre["fmean"] - y.mean()) < 1e-15 and np.allclose(pre["y"], yc.astype(np.float32), rtol=0, atol=0)
    # a duplicated column is aliased in lm(): the residual df stays n - rank
    pre2 = api.wbart_inputs(np.column_stack([X, X[:, 0]]), y, ntree=50, rm_const=True)
    assert abs(pre2["sigest"] - 1.0) < 1e-9
    # sigquant 0.99: qchisq(0.01, 3) = 0.1148318
    assert abs(api.wbart_inputs(X, y, sigquant=0.99)["lambda"] - 0.1148318 / 3.0) < 1e-7


def test_wbart_input_overrides():
    rng = np.random.default_rng(2)
    X = rng.random((20, 30))                                        # p >= n: sigest = sd(yc)
    y = rng.standard_normal(20)
    pre = api.wbart_inputs(X, y, ntree=10)
    assert abs(pre["sigest"] - np.std(y, ddof=1)) < 1e-12
    pre = api.wbart_inputs(X, y, lambda_=0.36, sigmaf=2.0, fmean=0.5, ntree=16, w=np.full(20, 2.0))
    assert abs(pre["sigest"] - 0.6) < 1e-15 and pre["lambda"] == 0.36
    assert abs(pre["tau"] - 0.5) < 1e-15 and pre["fmean"] == 0.5
    assert np.all(pre["wt"] == np.float32(0.25))
    assert np.allclose(pre["pgtab"][:3], [0.95, 0.95 / 4, 0.95 / 9], rtol=1e-7)


def test_pbart_inputs_known_answers():
    rng = np.random.default_rng(3)
    X = rng.random((40, 2))
    y = np.ones(40, int)
    y[7] = 0                                                        # mean 0.975
    pre = api.pbart_inputs(X, y, ntree=50)
    assert abs(pre["offset"] - 1.959963985) < 1e-8                  # qnorm(0.975)
    assert abs(pre["tau"] - 3.0 / (2.0 * math.sqrt(50))) < 1e-15
    assert pre["lambda"] == 1.0 and pre["nu"] == 3.0 and pre["sigest"] == 1.0
    assert np.all(pre["y"] == 0) and np.all(pre["wt"] == 1) and np.array_equal(pre["ybin"], y)
    assert api.pbart_inputs(X, y, binary_offset=0.25)["offset"] == 0.25


@pytest.mark.parametrize("case", ["p33", "ntree201", "numcut256", "n65537", "extra"])
def test_unsupported_settings_raise(case):
    rng = np.random.default_rng(4)
    if case == "p33":
        args = (rng.random((50, 33)), rng.random(50)), {}
    elif case == "ntree201":
        args = (rng.random((50, 2)), rng.random(50)), {"ntree": 201}
    elif case == "numcut256":
        args = (rng.random((50, 2)), rng.random(50)), {"xinfo": [np.linspace(0, 1, 256), np.linspace(0, 1, 10)]}
    elif case == "n65537":
        args = (rng.random((65537, 1)), rng.random(65537)), {}
    else:
        args = (rng.random((50, 2)), rng.random(50)), {"sparse": True}
    with pytest.raises(ValueError):
        api.wbart_inputs(*args[0], **args[1])


def test_objects_from_oracle_chain(oracle):
    rng = np.random.default_rng(5)
    X = rng.random((120, 3))
    y = np.sin(4 * X[:, 0]) + X[:, 1] + 0.3 * rng.standard_normal(120)
    pre = api.wbart_inputs(X, y, ntree=10, numcut=20)
    res = oracle.fit(pre, 10, 5, 7, 11)
    yh = oracle.yhat(res["trees"], pre["q"])
    fit = api.wbart_object(res, pre, yh)
    text, vc = oracle.tree_text(res["trees"], 3)
    assert fit["class"] == "wbart" and fit["treedraws"]["trees"] == text and np.array_equal(fit["varcount"], vc)
    assert np.array_equal(fit["yhat.train"], yh + pre["fmean"])
    assert np.array_equal(fit["yhat.train.mean"], res["trmean"].astype(np.float64) + pre["fmean"])
    assert np.array_equal(fit["sigma"], res["sigma"].astype(np.float64)) and fit["mu"] == pre["fmean"]
    assert np.allclose(fit["varprob"], 1 / 3) and fit["varprob"].shape == (7, 3)
    assert np.array_equal(fit["varcount.mean"], vc.mean(axis=0)) and np.array_equal(fit["moves"], res["stats"])
    # yhat.train.mean is the float running mean of the per-draw fits
    assert np.allclose(fit["yhat.train.mean"], fit["yhat.train"].mean(axis=0), atol=1e-5)

    yb = (y > np.median(y)).astype(int)
    pp = api.pbart_inputs(X, yb, ntree=10, numcut=20)
    res = oracle.fit(pp, 10, 5, 7, 12, probit=1)
    yh = oracle.yhat(res["trees"], pp["q"])
    fit = api.pbart_object(res, pp, yh)
    assert fit["class"] == "pbart" and "sigma" not in fit and "yhat.train.mean" not in fit
    assert np.array_equal(fit["yhat.train"], yh + pp["offset"])
    assert np.array_equal(fit["prob.train"], st.norm.cdf(yh + pp["offset"]))
    assert np.array_equal(fit["prob.train.mean"], fit["prob.train"].mean(axis=0))
    assert fit["binaryOffset"] == pp["offset"]
    # text from the library's writer equals the oracle's on these draws
    assert sampler.tree_text(res["trees"], 3)[0] == fit["treedraws"]["trees"]
