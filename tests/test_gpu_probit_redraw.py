"""The probit latent redraw spread over the device (the k_pr_* kernels: predicted bands, jump-ahead generation, band
decisions, per-block exit tables, the chained walk, the values) against the oracle's sequential redraw
(src/bart_kernel.cpp:543-547, rtnorm :282-292): a stream buffer shorter than the walk needs (slots past it stepped by
the walk itself), switches between the two redraws at run boundaries, the automatic choice on a larger batch,
hipGraph capture, and a chain of 65,536 rows."""
import sys

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd._lib import FbhError

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _check(b, chs, tag):
    for c, ch in enumerate(chs):
        s, e = b.state(c), ch.state()
        assert np.array_equal(s["rng"], e["rng"]), (tag, c)
        assert same_bits(s["latent"], e["latent"]), (tag, c)


@pytest.mark.parametrize("slots", [64, 1024])
def test_gpu_probit_redraw_short_stream_matches_oracle(oracle, slots):
    """With the stream buffer cut to `slots` slots, the rows past it are decided and valued from outputs the walk
    steps itself from the last checkpoint; nothing changes in the results."""
    probs = [W.small_problem(n, 3, 15, 300 + n, probit=True, ntree=8) for n in (700, 1333)]
    m, nskip, nd = 8, 4, 2
    chs = [oracle.chain(pb, m, nskip, nd, 31 + c, probit=1) for c, pb in enumerate(probs)]
    _say("short stream buffer", slots)
    with sampler.BatchSampler(probs, m, nskip, nd, [31, 32], probit=1, probit_redraw=2) as b:
        b.probit_redraw(2, slots)
        for k in range(3):
            b.run(2)
            for ch in chs:
                ch.sweeps(2)
            _check(b, chs, k)
        info = b.probit_info
        outs = [b.fetch(c) for c in range(2)]
    assert info["slots"] == slots and info["spot"] > 0
    for c, ch in enumerate(chs):
        assert same_bits(outs[c]["trees"].reshape(-1), ch.trees_out), c
        assert same_bits(outs[c]["trmean"], ch.trmean()), c


def test_gpu_probit_redraw_switches_between_runs(oracle):
    """Either redraw may follow the other at any run boundary (the generator state and latents carry over)."""
    probs = [W.small_problem(n, 4, 20, 400 + n, probit=True, ntree=10) for n in (513, 2048, 4100)]
    m, nskip, nd = 10, 6, 4
    chs = [oracle.chain(pb, m, nskip, nd, 41 + c, probit=1) for c, pb in enumerate(probs)]
    _say("redraw switches")
    with sampler.BatchSampler(probs, m, nskip, nd, [41, 42, 43], probit=1, probit_redraw=1,
                              sweeps_per_launch=2) as b:
        for k, mode in enumerate((1, 2, 1, 2, 2)):
            b.probit_redraw(mode)
            assert b.probit_info["mode"] == mode
            b.run(2)
            for ch in chs:
                ch.sweeps(2)
            _check(b, chs, k)
        outs = [b.fetch(c) for c in range(3)]
    for c, ch in enumerate(chs):
        assert same_bits(outs[c]["trees"].reshape(-1), ch.trees_out), c


def test_gpu_probit_redraw_automatic_choice_on_a_larger_batch(oracle):
    """Automatic: the sweep kernel's own redraw for more than 16 chains; forced to the device-wide one, a batch of 20
    chains of different lengths still matches the oracle chain by chain."""
    probs = [W.small_problem(200 + 97 * k, 3, 12, 500 + k, probit=True, ntree=6) for k in range(20)]
    m, nskip, nd = 6, 3, 2
    idx = [0, 7, 19]
    chs = {c: oracle.chain(probs[c], m, nskip, nd, 51 + c, probit=1) for c in idx}
    _say("automatic choice, 20 chains")
    with sampler.BatchSampler(probs, m, nskip, nd, [51 + c for c in range(20)], probit=1) as b:
        assert b.probit_info["mode"] == 1
        b.run(2)
        b.probit_redraw(2)
        b.run(3)
        for c, ch in chs.items():
            ch.sweeps(5)
            s, e = b.state(c), ch.state()
            assert np.array_equal(s["rng"], e["rng"]), c
            assert same_bits(s["latent"], e["latent"]), c


def test_gpu_probit_redraw_in_a_graph(oracle):
    """Runs captured as hipGraphs (the sweep kernel and the redraw kernels of every sweep), with eager state copies
    between them, against the oracle after every run."""
    pb = W.small_problem(1500, 3, 15, 611, probit=True, ntree=10)
    m, nskip, nd = 10, 4, 4
    ch = oracle.chain(pb, m, nskip, nd, 61, probit=1)
    _say("redraw in a graph")
    with sampler.BatchSampler([pb], m, nskip, nd, [61], probit=1, probit_redraw=2, graph=1) as b:
        for k in range(4):
            b.run(2)
            ch.sweeps(2)
            _check(b, [ch], k)
        out = b.fetch(0)
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)


def test_gpu_probit_redraw_rejects_bad_requests():
    pb = W.small_problem(300, 3, 10, 71, ntree=5)
    with sampler.BatchSampler([pb], 5, 1, 1, [1]) as b:
        assert b.probit_info["mode"] == 0
        with pytest.raises(FbhError):
            b.probit_redraw(2)
    pbp = W.small_problem(300, 3, 10, 72, probit=True, ntree=5)
    with sampler.BatchSampler([pbp], 5, 1, 1, [1], probit=1) as b:
        for mode, slots in ((3, 0), (-1, 0), (2, 100), (2, -64)):
            with pytest.raises(FbhError):
                b.probit_redraw(mode, slots)


def test_gpu_probit_redraw_large_chain_matches_oracle(oracle):
    """A probit chain of 65,536 rows (the largest rows per chain the workloads use; wider bands, more generator jump
    levels and block tables than the 24,576-row chain), a few sweeps, latents and generator state against the
    oracle."""
    pb = W.small_problem(65536, 4, 30, 811, probit=True, ntree=20)
    m, nskip, nd = 20, 2, 1
    ch = oracle.chain(pb, m, nskip, nd, 81, probit=1)
    _say("large probit chain, n = 65536")
    with sampler.BatchSampler([pb], m, nskip, nd, [81], probit=1, probit_redraw=2) as b:
        _check(b, [ch], "initial")
        for k in range(3):
            b.run(1)
            ch.sweeps(1)
            _check(b, [ch], k)
        info = b.probit_info
    _say("attempts decided outside the bands:", info["spot"])
    assert info["spot"] < 0.001 * 65536 * 4


@pytest.mark.parametrize("offset", [-2.5, 3.0])
@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_redraw_far_offsets_match_oracle(oracle, offset, redraw):
    """A binary offset far from the data (most attempts then truncate beyond the mean: rtnorm's exponential branch,
    :286-291, with its exact-image fallback) -- both redraws against the oracle."""
    pb = dict(W.small_problem(3000, 3, 15, 911, probit=True, ntree=10))
    pb["offset"] = float(np.float32(offset))
    m, nskip, nd = 10, 3, 2
    ch = oracle.chain(pb, m, nskip, nd, 91, probit=1)
    _say("far offset", offset, "redraw", redraw)
    with sampler.BatchSampler([pb], m, nskip, nd, [91], probit=1, probit_redraw=redraw) as b:
        _check(b, [ch], "initial")
        for k in range(nskip + nd):
            b.run(1)
            ch.sweeps(1)
            _check(b, [ch], k)
        out = b.fetch(0)
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)


@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_redraw_tiny_chains_match_oracle(oracle, redraw):
    """Chains shorter than one 64-row window (one partial block and group), beside a longer one, against the oracle."""
    probs = [W.small_problem(n, 2, 8, 950 + n, probit=True, ntree=4) for n in (7, 33, 63, 200)]
    m, nskip, nd = 4, 3, 2
    chs = [oracle.chain(pb, m, nskip, nd, 95 + c, probit=1) for c, pb in enumerate(probs)]
    _say("tiny chains, redraw", redraw)
    with sampler.BatchSampler(probs, m, nskip, nd, [95 + c for c in range(4)], probit=1, probit_redraw=redraw) as b:
        _check(b, chs, "initial")
        for k in range(nskip + nd):
            b.run(1)
            for ch in chs:
                ch.sweeps(1)
            _check(b, chs, k)


@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_fx64_and_node_cap_match_oracle(oracle, redraw):
    """Probit chains with fixed-point statistics and a small node cap (births refused at the cap), either redraw,
    against the oracle's FX64 twin: every output."""
    probs = [W.small_problem(n, 3, 20, 970 + n, probit=True, ntree=12) for n in (1500, 2600)]
    m, nskip, nd, cap = 12, 8, 4, 5
    _say("probit FX64, node cap", cap, "redraw", redraw)
    with sampler.BatchSampler(probs, m, nskip, nd, [97, 98], probit=1, stats_mode=1, node_cap=cap,
                              probit_redraw=redraw) as b:
        b.run(nskip + nd)
        outs = [b.fetch(c) for c in range(2)]
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, 97 + c, probit=1, stats_mode=1, node_cap=cap)
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (c, k)
    assert any(o["stats"][2] > 0 for o in outs)  # the cap was reached
