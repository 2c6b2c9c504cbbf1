"""GPU parity: the HIP sampler against the reference's trajectories and the oracle, bit for bit."""
import sys

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


GOLDEN = ["w37_p3", "w200_p3_extreme", "w1000_p16_ones", "w333_p1", "p300_probit", "w517_m1_deep",
          "w4096_p16_short", "n8_w800_cap"]


@pytest.mark.parametrize("name", GOLDEN)
def test_gpu_matches_reference_trajectory(golden_cases, name):
    g = golden_cases[name]
    _say("case", name)
    with sampler.BatchSampler([g["problem"]], g["m"], g["nskip"], g["ndpost"], [g["seed"]], probit=g["probit"],
                              node_cap=g["nodes"]) as b:
        b.run(g["nskip"] + g["ndpost"])
        out = b.fetch(0)
    e = g["expected"]
    assert same_bits(out["sigma"], e["sigma"])
    assert same_bits(out["trmean"], e["trmean"])
    assert np.array_equal(out["stats"], e["stats"])
    assert same_bits(out["trees"][:, :, :g["nodes"]], e["trees"])


def test_gpu_dropin_bart_fit_cpu(golden_cases):
    """The reference's own entry point (src/bart_kernel.h:63-66) on the device."""
    g = golden_cases["w1000_p16_ones"]
    out = sampler.bart_fit(g["problem"], g["m"], g["nskip"], g["ndpost"], g["seed"])
    e = g["expected"]
    for k in ("sigma", "trmean", "trees"):
        assert same_bits(out[k], e[k]), k
    assert np.array_equal(out["stats"], e["stats"])


def test_gpu_batch_of_uneven_chains_matches_oracle(oracle):
    """Chains of the 512-chain workload's shapes (uneven n and p) in one launch, each against the oracle."""
    idx = [0, 5, 17, 200, 333, 511]
    probs = [W.batch512_chain(c) for c in idx]
    seeds = [1 + c for c in idx]
    m, nskip, nd = 200, 5, 3
    _say("batch chains", idx)
    with sampler.BatchSampler(probs, m, nskip, nd, seeds) as b:
        b.run(nskip + nd)
        outs = [b.fetch(c) for c in range(len(idx))]
        yh = b.yhat(2)
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, seeds[c])
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (idx[c], k)
    # per-draw training fits of one chain, summed in tree order (src/glue.cpp:52-56)
    assert same_bits(yh, oracle.yhat(outs[2]["trees"], probs[2]["q"]))


def test_gpu_state_after_partial_run_and_resume(oracle):
    """Sampler state after k sweeps equals the oracle's (generator, sigma, running fit, leaf ids, trees); then the
    remaining sweeps run in further launches and the outputs still agree."""
    pb = W.small_problem(1500, 5, 30, 31, "uniform", ntree=30)
    m, nskip, nd = 30, 12, 6
    ch = oracle.chain(pb, m, nskip, nd, 11)
    with sampler.BatchSampler([pb], m, nskip, nd, [11], sweeps_per_launch=4) as b:
        b.run(7)
        ch.sweeps(7)
        s, e = b.state(0), ch.state()
        assert np.array_equal(s["rng"], e["rng"])
        assert s["sigma"] == e["sigma"]
        assert same_bits(s["af"], e["af"])
        assert np.array_equal(s["leaf"], e["leaf"])
        assert np.array_equal(s["trees"], e["trees"])
        b.run(nskip + nd - 7)
        ch.sweeps(nskip + nd - 7)
        out = b.fetch(0)
    assert same_bits(out["sigma"], ch.sigma_out)
    assert same_bits(out["trmean"], ch.trmean())
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)


@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_latents_match_oracle(oracle, redraw):
    """Latents, generator state and kept trees of a probit chain; redraw 1 walks the rows inside the sweep kernel,
    2 with the kernels spread over the device."""
    pb = W.small_problem(700, 3, 12, 41, probit=True, ntree=12)
    m, nskip, nd = 12, 10, 5
    ch = oracle.chain(pb, m, nskip, nd, 3, probit=1)
    with sampler.BatchSampler([pb], m, nskip, nd, [3], probit=1, probit_redraw=redraw) as b:
        assert b.probit_info["mode"] == redraw
        s0 = b.state(0)
        assert same_bits(s0["latent"], ch.state()["latent"])     # initial latents
        b.run(nskip + nd)
        ch.sweeps(nskip + nd)
        s = b.state(0)
        assert same_bits(s["latent"], ch.state()["latent"])
        out = b.fetch(0)
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)


def test_gpu_fx64_mode_matches_oracle(oracle):
    """Fixed-point statistics (order-independent sums) against the oracle's fixed-point twin."""
    probs = [W.small_problem(900, 4, 25, 51 + k, "uniform", ntree=20) for k in range(3)]
    m, nskip, nd = 20, 15, 5
    with sampler.BatchSampler(probs, m, nskip, nd, [7, 8, 9], stats_mode=1) as b:
        b.run(nskip + nd)
        outs = [b.fetch(c) for c in range(3)]
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, 7 + c, stats_mode=1)
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (c, k)


@pytest.mark.parametrize("stats_mode", [0, 1])
def test_gpu_streaming_variant_matches_oracle(oracle, stats_mode):
    """1024-row tiles with the running fit streamed from HBM (the variant large chains use), forced on small chains."""
    probs = [W.small_problem(n, 4, 20, 81 + n, "extreme", ntree=15) for n in (1000, 2600, 4133)]
    m, nskip, nd = 15, 10, 4
    _say("streaming variant, stats mode", stats_mode)
    with sampler.BatchSampler(probs, m, nskip, nd, [4, 5, 6], stats_mode=stats_mode, tile_rows=1024,
                              sweeps_per_launch=5) as b:
        b.run(nskip + nd)
        outs = [b.fetch(c) for c in range(3)]
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, 4 + c, stats_mode=stats_mode)
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (c, k)


def test_gpu_maxshape_chain_fx64_matches_oracle(oracle):
    """A chain of the max-shape workload (n = 65,536, p = 32, node cap 64, FX64), a few sweeps, against the oracle."""
    pb = W.maxshape64_chain(0)
    m, nskip, nd = 200, 2, 1
    _say("max-shape chain n", pb["q"].shape[1])
    with sampler.BatchSampler([pb], m, nskip, nd, [1], stats_mode=1, node_cap=64) as b:
        b.run(nskip + nd)
        out = b.fetch(0)
    e = oracle.fit(pb, m, nskip, nd, 1, stats_mode=1, node_cap=64)
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(out[k], e[k]), k


@pytest.mark.parametrize("tile_rows", [0, 1024])
@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_windows_match_oracle(oracle, tile_rows, redraw):
    """Latent redraw in windows of 64 rows: chains whose n is not a multiple of the window, resident and streamed
    running fit, latents and generator state after every launch against the oracle (both redraws)."""
    probs = [W.small_problem(n, 3, 15, 91 + n, probit=True, ntree=10) for n in (64, 130, 2999)]
    m, nskip, nd = 10, 6, 3
    chs = [oracle.chain(pb, m, nskip, nd, 21 + c, probit=1) for c, pb in enumerate(probs)]
    _say("probit windows, tile_rows", tile_rows, "redraw", redraw)
    with sampler.BatchSampler(probs, m, nskip, nd, [21, 22, 23], probit=1, tile_rows=tile_rows,
                              sweeps_per_launch=3, probit_redraw=redraw) as b:
        for k in range(3):
            b.run(3)
            for c, ch in enumerate(chs):
                ch.sweeps(3)
                s, e = b.state(c), ch.state()
                assert np.array_equal(s["rng"], e["rng"]), (c, k)
                assert same_bits(s["latent"], e["latent"]), (c, k)
        outs = [b.fetch(c) for c in range(3)]
    for c, ch in enumerate(chs):
        assert same_bits(outs[c]["trees"].reshape(-1), ch.trees_out), c
        assert same_bits(outs[c]["trmean"], ch.trmean()), c


@pytest.mark.parametrize("redraw", [1, 2])
def test_gpu_probit_workload_chain_matches_oracle(oracle, redraw):
    """The probit workload's chain (n = 24,576, p = 3, 50 trees), a few sweeps, with either redraw."""
    pb = W.workload_fpbart_n24576()
    m, nskip, nd = 50, 3, 2
    _say("probit workload chain, redraw", redraw)
    with sampler.BatchSampler([pb], m, nskip, nd, [1], probit=1, probit_redraw=redraw) as b:
        b.run(nskip + nd)
        s = b.state(0)
        out = b.fetch(0)
        info = b.probit_info
    if redraw == 2:
        # the bands hold nearly every row's path: attempts decided on the spot are rare
        print("attempts decided outside the bands:", info["spot"], "of", 24576 * (nskip + nd), "rows", flush=True)
        assert info["spot"] < 0.01 * 24576 * (nskip + nd)
    ch = oracle.chain(pb, m, nskip, nd, 1, probit=1)
    ch.sweeps(nskip + nd)
    e = ch.state()
    assert np.array_equal(s["rng"], e["rng"])
    assert same_bits(s["latent"], e["latent"])
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)
    assert same_bits(out["trmean"], ch.trmean())


_FULL = {}


@pytest.mark.parametrize("name", ["fwbart_n4096", "fpbart_n24576", "fpbart_n24576_in_kernel"])
def test_gpu_workload_chain_full_length_matches_oracle(oracle, name):
    """The single-chain workloads as configured: all 1,100 sweeps (nskip 1000, ndpost 100), every output bit for bit
    against the oracle (sigma trace, posterior-mean fit, the 100 kept draws' tree words, move counts); the probit
    chain with the automatic redraw (spread over the device) and with the one inside the sweep kernel."""
    redraw = 0
    if name == "fwbart_n4096":
        pb, m, probit = W.workload_fwbart_n4096(), 200, 0
    else:
        pb, m, probit = W.workload_fpbart_n24576(), 50, 1
        redraw = 1 if name.endswith("in_kernel") else 0
    nskip, nd = 1000, 100
    _say("full-length chain", name)
    with sampler.BatchSampler([pb], m, nskip, nd, [1], probit=probit, probit_redraw=redraw) as b:
        if probit:
            assert b.probit_info["mode"] == (redraw or 2)
        b.run(nskip + nd)
        out = b.fetch(0)
    key = name.replace("_in_kernel", "")
    if key not in _FULL:
        _FULL[key] = oracle.fit(pb, m, nskip, nd, 1, probit=probit)
    e = _FULL[key]
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(out[k], e[k]), k


def test_gpu_headline_batch_full_length_matches_oracle():
    """The headline configuration end to end: all 512 chains of the batched workload in one persistent launch per 20
    sweeps (the dense build, two chains per CU, serpentine packing, hipGraph capture, running fits in LDS), all 1,100
    sweeps; sixteen chains spread over the batch against the oracle, every output bit for bit."""
    from concurrent.futures import ProcessPoolExecutor
    import multiprocessing as mp
    import oracle_pool
    idx = list(range(0, 480, 32)) + [511]
    # the oracle's fits run in CPU worker processes (started before the GPU work) while the device samples
    pool = ProcessPoolExecutor(max_workers=min(len(idx), 8), mp_context=mp.get_context("spawn"))
    futs = [pool.submit(oracle_pool.fit_batch512_chain, (c, 200, 1000, 100)) for c in idx]
    probs = [W.batch512_chain(c) for c in range(512)]
    _say("headline batch: 512 chains, 1,100 sweeps")
    with sampler.BatchSampler(probs, 200, 1000, 100, [1 + c for c in range(512)], chains_per_cu=2, packing=1,
                              tile_rows=1024, resident=2, graph=1, sweeps_per_launch=20) as b:
        cfg = b.config
        assert (cfg["chains_per_cu"], cfg["packing"], cfg["tile_rows"], cfg["resident"], cfg["graph"]) == (2, 1, 1024,
                                                                                                           2, 1)
        b.run(1100)
        b.sync()
        outs = {c: b.fetch(c) for c in idx}
    _say("headline batch: device done; waiting for the oracle")
    exp = dict(f.result() for f in futs)
    pool.shutdown()
    for c in idx:
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], exp[c][k]), (c, k)
