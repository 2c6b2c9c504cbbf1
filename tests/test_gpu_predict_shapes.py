"""The batched predictor at the workloads' own predict shapes, against the oracle's restatement of predict_draw_cpp
(src/predict.cpp:49-73), bit for bit.

The rank-form kernel evaluates two rows per thread when a launch has at least 512 workgroups of 512 rows x draws;
these shapes take that instantiation (asserted through fbh_predict_last_launch):
  * 6 fits of the 512-chain workload's chains (200 trees, 100 kept draws each), all 600 draws on 5,237 new rows
    (ragged: the last workgroup of a draw holds 117 rows);
  * the probit workload's fit (50 trees, 100 kept draws) on its 24,576 training rows, and on a ragged prefix of them.
The fits run a shorter burn-in than the workloads (100 sweeps): the predict shape is the draws x rows x trees."""
import sys

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.predict import CompiledTrees, DeviceBuffer, DeviceRows, device_sync

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fit(probs, m, nskip, ndpost, probit=0):
    with sampler.BatchSampler(probs, m, nskip, ndpost, [1 + c for c in range(len(probs))], probit=probit,
                              keep_trees=1) as b:
        b.run(nskip + ndpost)
        return [b.fetch(c) for c in range(len(probs))]


def _compare_all_draws(oracle, ct, text, cuts, x, mu, got):
    oc = oracle.compile_trees(text, cuts)
    cols = [np.ascontiguousarray(x[:, v]) for v in range(x.shape[1])]
    for k in range(ct.ndpost):
        exp = oracle.predict_draw(oc, cols, k + 1, mu)
        assert same_bits(got[k], exp), k


def test_gpu_predict_batch512_shape_two_rows_per_thread(oracle):
    """600 draws (6 fits x 100) x 5,237 rows x 200 trees: every draw of every fit equals the oracle."""
    probs = [W.batch512_chain(c) for c in range(6)]
    _say("fitting 6 batch512 chains")
    outs = _fit(probs, 200, 100, 100)
    for c, pb in enumerate(probs):
        p = pb["q"].shape[0]
        text, _ = sampler.tree_text(outs[c]["trees"], p)
        mu = float(pb["fmean"])
        ct = CompiledTrees(text, pb["cutpoints"], mu=mu)
        assert ct.rank_form
        x = np.random.Generator(np.random.PCG64(4000 + c)).uniform(0.0, 1.0, size=(5237, p))
        x[::97, c % p] = pb["cutpoints"][c % p][7]      # rows exactly on a cutpoint go right
        rows = DeviceRows(x)
        d_draws = DeviceBuffer.from_array(np.arange(1, 101, dtype=np.int32))
        d_out = DeviceBuffer(8 * 100 * rows.n)
        rows.predict_device(ct, d_draws, 100, d_out)
        device_sync()
        assert ct.last_launch() == (1, 2)
        got = d_out.to_array(np.float64, (100, rows.n))
        _say("fit", c, "nodes per draw", ct.nnodes / ct.ndpost)
        _compare_all_draws(oracle, ct, text, pb["cutpoints"], x, mu, got)
        # the host entry point takes the same kernel and gives the same bits
        host = rows.predict(ct, np.arange(1, 101))
        assert ct.last_launch() == (1, 2)
        assert same_bits(host, got)
        for o in (d_out, d_draws, rows):
            o.close()
        ct.close()


def test_gpu_predict_probit_shape_two_rows_per_thread(oracle):
    """100 draws x 24,576 training rows x 50 trees (and a ragged prefix of 24,539 rows)."""
    pb = W.workload_fpbart_n24576()
    _say("fitting the probit workload chain")
    out = _fit([pb], 50, 100, 100, probit=1)[0]
    text, _ = sampler.tree_text(out["trees"], 3)
    mu = float(pb["offset"])
    ct = CompiledTrees(text, pb["cutpoints"], mu=mu)
    for n in (24576, 24539):
        x = np.ascontiguousarray(pb["x"][:n])
        got = ct.predict_draws(x, np.arange(1, 101))
        assert ct.last_launch() == (1, 2)
        _say("rows", n)
        _compare_all_draws(oracle, ct, text, pb["cutpoints"], x, mu, got)
    # a single draw is a small launch: one row per thread, same values
    one = ct.predict_draw(pb["x"], 37)
    assert ct.last_launch() == (1, 1)
    assert same_bits(one, ct.predict_draws(pb["x"], [37])[0])
    ct.close()


def test_gpu_predict_streams_share_a_handle(oracle):
    """Launches of one trees handle on the default stream and on a caller's non-blocking stream, alternating rows:
    the rank buffer is refilled and read in order, every result equals the oracle."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    probs = [W.batch512_chain(9)]
    out = _fit(probs, 200, 20, 10)[0]
    pb = probs[0]
    p = pb["q"].shape[0]
    text, _ = sampler.tree_text(out["trees"], p)
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.5)
    xs = [np.random.Generator(np.random.PCG64(77 + k)).uniform(0, 1, size=(3000 + 517 * k, p)) for k in range(3)]
    rows = [DeviceRows(x) for x in xs]
    d_draws = DeviceBuffer.from_array(np.arange(1, 11, dtype=np.int32))
    outs = [DeviceBuffer(8 * 10 * r.n) for r in rows]
    s = C.c_void_p()
    assert hip.hipStreamCreateWithFlags(C.byref(s), 1) == 0     # hipStreamNonBlocking, default priority
    from fastbart_amd._lib import check, lib
    for k in range(6):
        r = rows[k % 3]
        if k % 2:
            check(lib().fbh_predict_rows_device(ct.h, r.h, d_draws.ptr, 10, ct.mu, outs[k % 3].ptr, s),
                  "fbh_predict_rows_device")
        else:
            r.predict_device(ct, d_draws, 10, outs[k % 3])
    assert hip.hipStreamSynchronize(s) == 0
    device_sync()
    assert hip.hipStreamDestroy(s) == 0
    oc = oracle.compile_trees(text, pb["cutpoints"])
    for k in range(3):
        got = outs[k].to_array(np.float64, (10, rows[k].n))
        cols = [np.ascontiguousarray(xs[k][:, v]) for v in range(p)]
        for d in (1, 6, 10):
            assert same_bits(got[d - 1], oracle.predict_draw(oc, cols, d, 0.5)), (k, d)
    for o in outs + [d_draws] + rows:
        o.close()
    ct.close()
