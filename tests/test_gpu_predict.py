"""GPU predictor: compile_trees / predict_draw / batched draws against the oracle's restatement of src/predict.cpp,
and the invariant that re-predicting the training rows from the text reproduces yhat.train (the reference's
README.md:88).

Two device kernels serve a prediction: the rank form (sorted, NaN-free cutpoints: rows as 16-bit ranks, fixed-depth
walks) and the general form (any cutpoint table, any draw size). Both are compared with the oracle bit for bit."""
import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.predict import CompiledTrees, DeviceBuffer, DeviceRows, device_sync

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


def _new_rows(pb, n=777, seed=0):
    """Rows outside the training range, NaNs, and values exactly on cutpoints (x >= cut is true there)."""
    rng = np.random.default_rng(seed)
    x = rng.uniform(-0.1, 1.1, size=(n, 4))
    x[::50, 2] = np.nan                          # NaN goes left (x >= cut is false)
    for v in range(4):
        xi = np.asarray(pb["cutpoints"][v])
        if len(xi):
            k = rng.integers(0, len(xi), size=n // 7)
            x[3::7, v][:len(k)] = xi[k]
    x[5, :] = -np.inf
    x[6, :] = np.inf
    return x


def _check_against_oracle(oracle, ct, oc, text, xs, draws, mu):
    for x in xs:
        cols = [np.ascontiguousarray(x[:, v]) for v in range(x.shape[1])]
        for m_draw in draws:
            got = ct.predict_draw(x, m_draw)
            exp = oracle.predict_draw(oc, cols, m_draw, mu)
            assert same_bits(got, exp), m_draw


def test_gpu_predict_draw_matches_oracle(oracle, fitted):
    pb, out, yh, text = fitted
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.25)
    assert ct.rank_form
    oc = oracle.compile_trees(text, pb["cutpoints"])
    xnew = _new_rows(pb)
    _check_against_oracle(oracle, ct, oc, text, (pb["x"], xnew), (1, 7, 12), 0.25)
    allx = ct.predict_draws(xnew, np.arange(1, 13))
    for m_draw in range(1, 13):
        assert same_bits(allx[m_draw - 1], ct.predict_draw(xnew, m_draw))
    rows = DeviceRows(xnew)
    assert same_bits(rows.predict(ct, [3, 3, 9]), allx[[2, 2, 8]])


def test_gpu_predict_general_form_matches_oracle(oracle, fitted):
    """A cutpoint table that is not sorted (each variable's cutpoints reversed, the tree text unchanged) takes the
    general kernel; a table with ties stays in the rank form. Both equal the oracle."""
    pb, out, yh, text = fitted
    xnew = _new_rows(pb, 600, seed=3)
    rev = [np.asarray(c)[::-1].copy() for c in pb["cutpoints"]]
    ct = CompiledTrees(text, rev, mu=-1.5)
    assert not ct.rank_form
    oc = oracle.compile_trees(text, rev)
    _check_against_oracle(oracle, ct, oc, text, (pb["x"], xnew), (1, 5, 12), -1.5)
    ties = [np.sort(np.round(np.asarray(c) * 8) / 8) for c in pb["cutpoints"]]
    ct = CompiledTrees(text, ties, mu=0.0)
    assert ct.rank_form
    oc = oracle.compile_trees(text, ties)
    _check_against_oracle(oracle, ct, oc, text, (xnew, np.round(xnew * 8) / 8), (2, 11), 0.0)


def _full_tree_text(ndpost, ntree, depth, p, ncut, seed):
    """Complete binary trees of the given depth in the treedraws text format (pre-order ids 1, 2id, 2id+1)."""
    rng = np.random.default_rng(seed)
    lines = [f"{ndpost} {ntree} {p}"]
    for _ in range(ndpost * ntree):
        nodes = []

        def rec(i, d):
            if d == depth:
                nodes.append(f"{i} 0 0 {rng.normal():.10g}")
                return
            nodes.append(f"{i} {rng.integers(0, p)} {rng.integers(0, ncut)} 0")
            rec(2 * i, d + 1)
            rec(2 * i + 1, d + 1)

        rec(1, 0)
        lines.append(str(len(nodes)))
        lines.extend(nodes)
    return "\n".join(lines) + "\n"


def test_gpu_predict_large_draws_take_the_general_kernel(oracle):
    """200 complete trees of depth 5 (12,600 nodes per draw) exceed the rank kernel's LDS: the general kernel runs
    and still equals the oracle; depth-5 trees of 40 per draw fit and take the rank kernel."""
    p, ncut = 3, 30
    rng = np.random.default_rng(9)
    cuts = [np.sort(rng.uniform(0, 1, ncut)) for _ in range(p)]
    x = rng.uniform(0, 1, size=(700, p))
    cols = [np.ascontiguousarray(x[:, v]) for v in range(p)]
    for ntree in (200, 40):
        text = _full_tree_text(3, ntree, 5, p, ncut, seed=ntree)
        ct = CompiledTrees(text, cuts, mu=0.5)
        oc = oracle.compile_trees(text, cuts)
        got = ct.predict_draws(x, [1, 2, 3])
        for m_draw in (1, 2, 3):
            assert same_bits(got[m_draw - 1], oracle.predict_draw(oc, cols, m_draw, 0.5)), (ntree, m_draw)


def test_gpu_predict_device_entry_point(fitted):
    """fbh_predict_rows_device: device draw list and output, bit-identical to the host entry point; a draw outside
    1..ndpost gives a NaN row and leaves the others intact."""
    pb, out, yh, text = fitted
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.125)
    xnew = _new_rows(pb, 1031, seed=5)
    rows = DeviceRows(xnew)
    draws = np.array([12, 1, 0, 7, 13, 7], np.int32)
    d_draws = DeviceBuffer.from_array(draws)
    d_out = DeviceBuffer(8 * len(draws) * rows.n)
    rows.predict_device(ct, d_draws, len(draws), d_out)
    device_sync()
    got = d_out.to_array(np.float64, (len(draws), rows.n))
    ok = [0, 1, 3, 5]
    assert same_bits(got[ok], rows.predict(ct, draws[ok]))
    assert np.isnan(got[2]).all() and np.isnan(got[4]).all()


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
