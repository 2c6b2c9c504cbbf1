"""Row split (one chain over several workgroups, FX64 statistics): every member runs the same control steps and the
members add their fixed-point pass totals after each pass. Fixed-point sums are order-free, so a split chain must equal
the one-workgroup FX64 chain and the oracle's FX64 twin bit for bit: trees, sigma trace, trmean, move counts, per-draw
training fits and the final chain state."""
import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd._lib import FbhError

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[row split]", *a, flush=True)


@pytest.mark.parametrize("split", [2, 3, 8])
def test_gpu_row_split_matches_oracle_fx64(oracle, split):
    probs = [W.small_problem(n, 4, 25, 61 + n, "uniform", ntree=20) for n in (1000, 1555, 2999)]
    m, nskip, nd = 20, 12, 6
    _say("split", split)
    with sampler.BatchSampler(probs, m, nskip, nd, [3, 4, 5], stats_mode=1, row_split=split,
                              sweeps_per_launch=7) as b:
        b.run(nskip + nd)
        outs = [b.fetch(c) for c in range(3)]
        yh = [b.yhat(c) for c in range(3)]
    with sampler.BatchSampler(probs, m, nskip, nd, [3, 4, 5], stats_mode=1) as b:
        b.run(nskip + nd)
        ref = [b.fetch(c) for c in range(3)]
        yref = [b.yhat(c) for c in range(3)]
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, 3 + c, stats_mode=1)
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (c, k)
            assert same_bits(outs[c][k], ref[c][k]), (c, k)
        assert same_bits(yh[c], yref[c]), c


def test_gpu_row_split_state_across_launches():
    """State after partial runs (launch boundaries inside the fit) equals the unsplit chain's: generator, sigma, running
    fit, leaf ids and tree words."""
    probs = [W.small_problem(n, 5, 30, 71 + n, "extreme", ntree=15) for n in (777, 2048)]
    m, nskip, nd = 15, 8, 4
    _say("state across launches")
    with sampler.BatchSampler(probs, m, nskip, nd, [9, 10], stats_mode=1, row_split=4, sweeps_per_launch=3) as a, \
            sampler.BatchSampler(probs, m, nskip, nd, [9, 10], stats_mode=1, sweeps_per_launch=3) as b:
        for _ in range(4):
            a.run(3)
            b.run(3)
            for c in range(2):
                sa, sb = a.state(c), b.state(c)
                for k in ("rng", "af", "leaf", "trees"):
                    assert same_bits(sa[k], sb[k]), k
                assert sa["sigma"] == sb["sigma"]


def test_gpu_row_split_maxshape_chain():
    """A chain of the max-shape workload (n = 65,536, p = 32, node cap 64) split over 8 workgroups, a few sweeps,
    against the same chain on one workgroup."""
    pb = W.maxshape64_chain(1)
    m, nskip, nd = 200, 2, 1
    _say("max-shape chain split 8")
    outs = []
    for split in (8, 0):
        with sampler.BatchSampler([pb], m, nskip, nd, [2], stats_mode=1, node_cap=64, row_split=split) as b:
            b.run(nskip + nd)
            outs.append(b.fetch(0))
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(outs[0][k], outs[1][k]), k


def test_gpu_row_split_rejects_what_it_cannot_run():
    pb = W.small_problem(1000, 4, 25, 5, "uniform", ntree=10)
    with pytest.raises(FbhError):                       # REF16 order needs one workgroup per lane sequence
        sampler.BatchSampler([pb], 10, 2, 2, [1], stats_mode=0, row_split=2)
    with pytest.raises(FbhError):                       # fewer than 64 rows per member
        sampler.BatchSampler([pb], 10, 2, 2, [1], stats_mode=1, row_split=16)
    with pytest.raises(FbhError):                       # more workgroups than fit on the device at once
        sampler.BatchSampler([pb] * 1100, 10, 2, 2, list(range(1100)), stats_mode=1, row_split=2)


def test_gpu_row_split_automatic():
    """row_split = -1: up to one workgroup per CU for FX64 batches, with at least 3,072 rows per member (here the
    shorter chain's 7,000 rows allow 2); none for REF16, nor for chains shorter than 6,144 rows."""
    probs = [W.small_problem(n, 4, 25, 31 + n, "uniform", ntree=12) for n in (13000, 7000)]
    m, nskip, nd = 12, 6, 4
    _say("automatic split")
    outs = {}
    for split in (-1, 0):
        with sampler.BatchSampler(probs, m, nskip, nd, [5, 6], stats_mode=1, row_split=split) as b:
            if split == -1:
                assert b.row_split == 2
            b.run(nskip + nd)
            outs[split] = [b.fetch(c) for c in range(2)]
    for c in range(2):
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[-1][c][k], outs[0][c][k]), (c, k)
    with sampler.BatchSampler(probs, m, nskip, nd, [5, 6], stats_mode=0, row_split=-1) as b:
        assert b.row_split == 1
    short = [W.small_problem(4096, 4, 25, 37, "uniform", ntree=12)]
    with sampler.BatchSampler(short, m, nskip, nd, [5], stats_mode=1, row_split=-1) as b:
        assert b.row_split == 1


def test_gpu_row_split_with_launch_configuration(oracle):
    """A row-split batch under the launch configuration: graph capture, the autotuner's trial launches (only
    configurations whose workgroups are all resident at once), and a caller's packing mid-run, against the oracle's
    FX64 twin."""
    probs = [W.small_problem(n, 4, 25, 77 + n, "uniform", ntree=12) for n in (900, 1500, 2100)]
    m, nskip, nd = 12, 18, 6
    seeds = [11, 12, 13]
    _say("split with autotune, graph and a caller's order")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, stats_mode=1, row_split=3, graph=1,
                              sweeps_per_launch=4) as b:
        t = b.autotune(1)
        assert b.it == t["sweeps"]
        b.set_order([2, 0, 1])
        b.run(5)
        b.configure(1, 0, 1, 0)
        b.run(nskip + nd - b.it)
        b.sync()
        for c, pb in enumerate(probs):
            e = oracle.fit(pb, m, nskip, nd, seeds[c], stats_mode=1)
            out = b.fetch(c)
            for k in ("sigma", "trmean", "trees", "stats"):
                assert same_bits(out[k], e[k]), (c, k)


@pytest.mark.parametrize("n,split", [(24576, 4), (5000, 3), (24576, 16)])
def test_gpu_row_split_probit_fx64_matches_oracle(oracle, n, split):
    """A probit chain with FX64 statistics split over workgroups (each member 64-row multiples of the chain's rows,
    the pass totals exchanged after every pass) and its latents redrawn by the device-wide kernels over all members'
    rows: latents and generator state after every sweep, and every output, against the oracle's one-chain FX64 twin."""
    pb = W.small_problem(n, 3, 25, 1200 + n, probit=True, ntree=20)
    m, nskip, nd = 20, 3, 2
    ch = oracle.chain(pb, m, nskip, nd, 7, probit=1, stats_mode=1)
    print("probit row split", n, split, flush=True)
    with sampler.BatchSampler([pb], m, nskip, nd, [7], probit=1, stats_mode=1, row_split=split) as b:
        assert b.row_split == split and b.probit_info["mode"] == 2
        s0, e0 = b.state(0), ch.state()
        assert same_bits(s0["latent"], e0["latent"])
        for k in range(nskip + nd):
            b.run(1)
            ch.sweeps(1)
            s, e = b.state(0), ch.state()
            assert np.array_equal(s["rng"], e["rng"]), k
            assert same_bits(s["latent"], e["latent"]), k
        out = b.fetch(0)
        with pytest.raises(FbhError):                   # the in-kernel redraw sees one member's rows only
            b.probit_redraw(1)
    assert same_bits(out["trees"].reshape(-1), ch.trees_out)
    assert same_bits(out["trmean"], ch.trmean())


def test_gpu_row_split_probit_rejects_what_it_cannot_run():
    pb = W.small_problem(3000, 3, 20, 1301, probit=True, ntree=5)
    with pytest.raises(FbhError):                       # REF16
        sampler.BatchSampler([pb], 5, 1, 1, [1], probit=1, stats_mode=0, row_split=2)
    with pytest.raises(FbhError):                       # the in-kernel redraw asked for
        sampler.BatchSampler([pb], 5, 1, 1, [1], probit=1, stats_mode=1, row_split=2, probit_redraw=1)
    with pytest.raises(FbhError):                       # 200 rows in 128-row members: the third would be empty
        sampler.BatchSampler([W.small_problem(200, 3, 20, 1302, probit=True, ntree=5)], 5, 1, 1, [1], probit=1,
                             stats_mode=1, row_split=3)
