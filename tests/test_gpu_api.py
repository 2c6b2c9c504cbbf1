"""fwbart() / fpbart() / fwbart_batch() end to end on the device: sampler outputs bit-equal to the oracle on the same
preprocessed inputs, the object's fields built from them, and compile_trees + predict_draw on the training rows:
bit-equal to the oracle's predict_draw on the same text, and equal to yhat.train up to the text's 10 significant
digits per leaf value (pwbart(x.train)[m, ] reproduces yhat.train[m, ] only to that precision, in R too)."""
import sys

import numpy as np
import pytest
from scipy import stats as st

from conftest import same_bits
from fastbart_amd import api, predict

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check_predict(oracle, fit, X, draws, ntree):
    ct = predict.compile_trees(fit)
    oc = oracle.compile_trees(fit["treedraws"]["trees"], fit["treedraws"]["cutpoints"])
    mu = ct.mu
    cols = [np.ascontiguousarray(X[:, v - 1]) for v in fit["rm.const"]]
    try:
        for m in draws:
            f = predict.predict_draw(ct, X, m)
            assert np.array_equal(f, oracle.predict_draw(oc, cols, m, mu)), m
            tol = ntree * 5e-9 * max(1.0, float(np.abs(fit["yhat.train"][m - 1]).max()))
            assert np.max(np.abs(f - fit["yhat.train"][m - 1])) <= tol, m
    finally:
        ct.close()


def _data(seed, n=300, p=4, const_col=False):
    rng = np.random.default_rng(seed)
    X = rng.random((n, p))
    if const_col:
        X[:, 1] = 2.5
    y = 3 * np.sin(3 * X[:, 0]) + 2 * (X[:, 2] > 0.4) + 0.5 * rng.standard_normal(n)
    return X, y, rng.uniform(0.5, 2.0, n)


@pytest.mark.parametrize("const_col", [False, True])
def test_fwbart_matches_oracle_and_predicts_its_training_fit(oracle, const_col):
    X, y, w = _data(1, const_col=const_col)
    _say("fwbart const_col", const_col)
    kw = dict(w=w, ntree=20, numcut=30)
    fit = api.fwbart(X, y, ndpost=25, nskip=15, seed=7, **kw)
    pre = api.wbart_inputs(X, y, **kw)
    e = oracle.fit(pre, 20, 15, 25, 7)
    assert same_bits(fit["sigma"].astype(np.float32), e["sigma"])
    assert np.array_equal(fit["yhat.train.mean"], e["trmean"].astype(np.float64) + pre["fmean"])
    assert np.array_equal(fit["yhat.train"], oracle.yhat(e["trees"], pre["q"]) + pre["fmean"])
    assert fit["treedraws"]["trees"] == oracle.tree_text(e["trees"], pre["q"].shape[0])[0]
    assert np.array_equal(fit["moves"], e["stats"])
    assert np.array_equal(fit["rm.const"], [1, 3, 4] if const_col else [1, 2, 3, 4])
    _check_predict(oracle, fit, X, (1, 13, 25), 20)


def test_fpbart_matches_oracle_and_predicts_its_training_fit(oracle):
    X, y, _ = _data(2, n=400, p=3)
    yb = (y > np.median(y)).astype(int)
    _say("fpbart n400")
    fit = api.fpbart(X, yb, ndpost=20, nskip=10, seed=9, ntree=30, numcut=25)
    pre = api.pbart_inputs(X, yb, ntree=30, numcut=25)
    e = oracle.fit(pre, 30, 10, 20, 9, probit=1)
    yh = oracle.yhat(e["trees"], pre["q"])
    assert np.array_equal(fit["yhat.train"], yh + pre["offset"])
    assert np.array_equal(fit["prob.train"], st.norm.cdf(yh + pre["offset"]))
    assert fit["treedraws"]["trees"] == oracle.tree_text(e["trees"], 3)[0]
    assert np.array_equal(fit["moves"], e["stats"])
    _check_predict(oracle, fit, X, (1, 20), 30)


def test_fwbart_batch_equals_fwbart_per_data_set():
    sets = [_data(10 + i, n=150 + 97 * i, p=3 + i) for i in range(3)]
    _say("fwbart_batch 3 sets")
    fits = api.fwbart_batch(sets, ndpost=10, nskip=10, seeds=[3, 4, 5], ntree=15, numcut=20)
    for i, (X, y, w) in enumerate(sets):
        one = api.fwbart(X, y, w=w, ndpost=10, nskip=10, seed=3 + i, ntree=15, numcut=20)
        for k in ("sigma", "yhat.train", "yhat.train.mean", "varcount", "moves"):
            assert np.array_equal(fits[i][k], one[k]), (i, k)
        assert fits[i]["treedraws"]["trees"] == one["treedraws"]["trees"]


def test_fpbart_batch_equals_fpbart_per_data_set():
    sets = []
    for i in range(2):
        X, y, _ = _data(20 + i, n=200 + 150 * i, p=3)
        sets.append((X, (y > np.median(y)).astype(int)))
    _say("fpbart_batch 2 sets")
    fits = api.fpbart_batch(sets, ndpost=8, nskip=6, seeds=[5, 6], ntree=12, numcut=20)
    for i, (X, yb) in enumerate(sets):
        one = api.fpbart(X, yb, ndpost=8, nskip=6, seed=5 + i, ntree=12, numcut=20)
        for k in ("yhat.train", "prob.train", "varcount", "moves"):
            assert np.array_equal(fits[i][k], one[k]), (i, k)
        assert fits[i]["treedraws"]["trees"] == one["treedraws"]["trees"]
