"""GPU predictor: compile_trees / predict_draw / batched draws against the oracle's restatement of src/predict.cpp,
and the invariant that re-predicting the training rows from the text reproduces yhat.train (README.md:88)."""
import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.predict import CompiledTrees, DeviceRows

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fitted():
    pb = W.small_problem(1200, 4, 20, 61, "uniform", ntree=25)
    with sampler.BatchSampler([pb], 25, 20, 12, [5]) as b:
        b.run(32)
        out = b.fetch(0)
        yh = b.yhat(0)
    text, vc = sampler.tree_text(out["trees"], 4)
    return pb, out, yh, text


def test_gpu_predict_draw_matches_oracle(oracle, fitted):
    pb, out, yh, text = fitted
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.25)
    oc = oracle.compile_trees(text, pb["cutpoints"])
    rng = np.random.default_rng(0)
    xnew = rng.uniform(-0.1, 1.1, size=(777, 4))
    xnew[::50, 2] = np.nan                      # NaN goes left (x >= cut is false)
    for x in (pb["x"], xnew):
        cols = [np.ascontiguousarray(x[:, v]) for v in range(4)]
        for m_draw in (1, 7, 12):
            got = ct.predict_draw(x, m_draw)
            exp = oracle.predict_draw(oc, cols, m_draw, 0.25)
            assert same_bits(got, exp), m_draw
    allx = ct.predict_draws(xnew, np.arange(1, 13))
    for m_draw in range(1, 13):
        assert same_bits(allx[m_draw - 1], ct.predict_draw(xnew, m_draw))
    rows = DeviceRows(xnew)
    assert same_bits(rows.predict(ct, [3, 3, 9]), allx[[2, 2, 8]])


def test_gpu_predict_reproduces_training_fits(fitted):
    pb, out, yh, text = fitted
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.0)
    pr = ct.predict_draws(pb["x"], np.arange(1, 13))
    assert np.max(np.abs(pr - yh)) < 1e-9 * max(1.0, np.max(np.abs(yh)))


def test_gpu_predict_rejects_bad_draw(fitted):
    pb, out, yh, text = fitted
    ct = CompiledTrees(text, pb["cutpoints"])
    from fastbart_amd._lib import FbhError
    with pytest.raises(FbhError):
        ct.predict_draw(pb["x"], 13)
    with pytest.raises(FbhError):
        ct.predict_draw(pb["x"][:, :3], 1)
