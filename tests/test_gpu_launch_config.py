"""Launch configuration of the sweep kernel: the workgroup -> chain packing, the tile rows, the build (one chain per
CU at 168 VGPRs or two at 128), graph capture and the autotuner that picks among them. Chains are independent and the state between launches is complete in device memory,
so none of these may change a result: every configuration, and a switch between configurations at any launch
boundary, must reproduce the oracle's chains bit for bit."""
import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd._lib import FbhError

pytestmark = pytest.mark.gpu

NS = (300, 1700, 901, 2500, 1203, 640)


def _say(*a):
    print("[launch config]", *a, flush=True)


def _probs(probit=False, ntree=12):
    return [W.small_problem(n, 4, 20, 510 + n, "uniform", probit=probit, ntree=ntree) for n in NS]


def _check_oracle(oracle, b, probs, m, nskip, nd, seeds, probit=0):
    for c, pb in enumerate(probs):
        e = oracle.fit(pb, m, nskip, nd, seeds[c], probit=probit)
        out = b.fetch(c)
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(out[k], e[k]), (c, k)


@pytest.mark.parametrize("packing,graph,per_cu", [(1, 1, 2), (2, 0, 1), (0, 1, 0)])
def test_gpu_packing_and_graph_match_oracle(oracle, packing, graph, per_cu):
    probs = _probs()
    m, nskip, nd = 12, 14, 6
    seeds = [31 + c for c in range(len(probs))]
    _say("packing", packing, "graph", graph, "chains per CU", per_cu)
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, packing=packing, graph=graph, sweeps_per_launch=4,
                              chains_per_cu=per_cu) as b:
        assert b.config["packing"] == packing and b.config["graph"] == graph
        b.run(7)
        b.run(nskip + nd - 7)
        b.sync()
        avg, nl, _ = b.timing()
        assert nl == 2 + 4 and avg > 0  # 7 sweeps in 2 launches, 13 in 4 (a graph is timed as a whole)
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)


@pytest.mark.parametrize("per_cu", [1, 2])
def test_gpu_trees_past_six_leaves_match_oracle(oracle, per_cu):
    """Two trees on 4,000 rows grow past six leaves. The one-chain-per-CU build tags a row with one bit at its leaf's
    rank while the tree has at most six leaves and with the rank itself beyond that; the dense build always with the
    rank. Every output against the oracle in both builds, with kept trees on both sides of six leaves."""
    pb = W.small_problem(4000, 5, 40, 977, ntree=2)
    m, nskip, nd = 2, 0, 60  # kept from the first sweep: 2 to 17 leaves
    _say("trees past six leaves, chains per CU", per_cu)
    with sampler.BatchSampler([pb], m, nskip, nd, [7], chains_per_cu=per_cu, sweeps_per_launch=10) as b:
        b.run(nskip + nd)
        out = b.fetch(0)
    e = oracle.fit(pb, m, nskip, nd, 7)
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(out[k], e[k]), k
    w = np.asarray(out["trees"]).view(np.uint64).reshape(-1, 64)
    leaves = (((w & 255) != 255) & (((w >> 16) & 255) == 0) & (((w >> 24) & 255) == 0)).sum(axis=1)
    assert leaves.max() > 6 and leaves.min() <= 6, (leaves.min(), leaves.max())


def test_gpu_probit_packing_and_graph_match_oracle(oracle):
    probs = _probs(probit=True, ntree=8)
    m, nskip, nd = 8, 9, 5
    seeds = [41 + c for c in range(len(probs))]
    _say("probit, packing 1, graph")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, probit=1, packing=1, graph=1, sweeps_per_launch=3,
                              chains_per_cu=2) as b:
        b.run(nskip + nd)
        b.sync()
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds, probit=1)


def test_gpu_configuration_switched_between_launches(oracle):
    """Packing, tile kind and build changed at every launch boundary: the state carried between launches is the same
    in every configuration."""
    probs = _probs()
    m, nskip, nd = 12, 18, 6
    seeds = [51 + c for c in range(len(probs))]
    # (packing, tile_rows, resident, chains_per_cu) asked -> (tile_rows, resident) in effect
    cycle = [((0, 1024, 1, 1), (1024, 1)), ((1, 1536, 0, 2), (1536, 2)), ((2, 1024, 2, 2), (1024, 2)),
             ((2, 1536, 2, 1), (1536, 2)), ((1, 1024, 0, 0), (1024, 1)), ((0, 0, 2, 2), (1024, 2)),
             ((1, 0, 0, 1), (1536, 2)), ((2, 1024, 2, 1), (1024, 2))]
    _say("switched configurations")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=3) as b:
        cfg = b.config
        # automatic: six workgroups (fewer than the CUs) -> one chain per CU, 1536-row tiles with the fit in LDS
        assert (cfg["chains_per_cu"], cfg["tile_rows"], cfg["resident"]) == (1, 1536, 2)
        prev = 1
        for (pk, tr, res, cp), (tr_eff, res_eff) in cycle:
            b.configure(pk, tr, cp, res)
            got = b.config
            assert (got["packing"], got["tile_rows"], got["resident"], got["chains_per_cu"]) == \
                (pk, tr_eff, res_eff, cp or prev)
            prev = got["chains_per_cu"]
            b.run(3)
        b.sync()
        assert b.it == nskip + nd
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)


def test_gpu_caller_order_matches_oracle(oracle):
    probs = _probs()
    m, nskip, nd = 12, 10, 5
    seeds = [71 + c for c in range(len(probs))]
    _say("caller's order")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=3) as b:
        b.set_order([3, 0, 5, 1, 4, 2])
        assert b.config["packing"] == -1
        b.run(8)
        b.set_order([5, 4, 3, 2, 1, 0])
        b.run(nskip + nd - 8)
        b.sync()
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)
        for bad in ([0, 1, 2, 3, 4, 4], [0, 1, 2, 3, 4, 6], [-1, 1, 2, 3, 4, 5]):
            with pytest.raises(FbhError):
                b.set_order(bad)


def test_gpu_autotune_runs_real_sweeps_and_keeps_results(oracle):
    probs = _probs()
    m, nskip, nd = 12, 16, 6
    seeds = [61 + c for c in range(len(probs))]
    _say("autotune")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=5, graph=1) as b:
        t = b.autotune(1)
        # stage 1: the three tile kinds at one chain per CU (six workgroups, fewer than the CUs); stage 2: packings
        c1, c2 = t["candidates"][:3], t["candidates"][3:]
        assert len(c1) == 3 and len(c2) == 3
        assert {(c["tile_rows"], c["resident"]) for c in c1} == {(1024, 1), (1024, 2), (1536, 2)}
        assert all(c["chains_per_cu"] == 1 for c in c1)
        assert [c["packing"] for c in c2] == [0, 1, 2]
        assert t["sweeps"] == 13 and b.it == 13
        assert all(c["ms_per_sweep"] > 0 for c in t["candidates"])
        w1 = min(c1, key=lambda c: c["ms_per_sweep"])
        assert all((c["tile_rows"], c["resident"]) == (w1["tile_rows"], w1["resident"]) for c in c2)
        best = min(c2, key=lambda c: c["ms_per_sweep"])
        assert (t["packing"], t["tile_rows"], t["resident"], t["chains_per_cu"]) == \
            (best["packing"], best["tile_rows"], best["resident"], 1)
        cfg = b.config
        assert (cfg["packing"], cfg["tile_rows"], cfg["resident"], cfg["chains_per_cu"]) == \
            (t["packing"], t["tile_rows"], t["resident"], 1)
        assert cfg["graph"] == 1  # graph capture is left as it was
        b.run(nskip + nd - 13)
        b.sync()
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)


def test_gpu_autotune_cache(oracle, tmp_path):
    """The autotuner cache: a new batch shape is tuned and its choice appended; a batch of the same shape (other
    seeds) takes the stored choice without a sweep; the last entry for a shape wins; an entry the batch cannot run is
    tuned afresh; another shape is a miss. Every run still matches the oracle."""
    probs = _probs()
    m, nskip, nd = 12, 16, 6
    cache = tmp_path / "tune.txt"
    _say("autotune cache: miss")
    seeds = [71 + c for c in range(len(probs))]
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=5) as b:
        t = b.autotune(1, cache=str(cache))
        assert not t["from_cache"] and t["sweeps"] == 13 and b.it == 13
        lines = cache.read_text().splitlines()
        assert len(lines) == 1
        key, vals = lines[0].split("\t")
        f = vals.split()
        assert [int(x) for x in f[:4]] == [t["packing"], t["tile_rows"], t["resident"], t["chains_per_cu"]]
        assert float(f[4]) == pytest.approx(min(c["ms_per_sweep"] for c in t["candidates"][3:]), rel=1e-5)
        assert "chains=6" in key and "ntree=12" in key and "stats=ref16" in key
        b.run(nskip + nd - 13)
        b.sync()
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)
    _say("autotune cache: hit")
    seeds = [81 + c for c in range(len(probs))]
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=5) as b:
        t2 = b.autotune(1, cache=str(cache))
        assert t2["from_cache"] and t2["sweeps"] == 0 and b.it == 0 and t2["candidates"] == []
        assert all(t2[k] == t[k] for k in ("packing", "tile_rows", "resident", "chains_per_cu"))
        cfg = b.config
        assert all(cfg[k] == t[k] for k in ("packing", "tile_rows", "resident", "chains_per_cu"))
        b.run(nskip + nd)
        b.sync()
        _check_oracle(oracle, b, probs, m, nskip, nd, seeds)
    _say("autotune cache: last entry wins, malformed lines skipped")
    other = (t["packing"] + 1) % 3
    with open(cache, "a") as fh:
        fh.write("no tab on this line\n")
        fh.write(f"{key}\tnot numbers\n")
        fh.write(f"{key}\t{other} 1024 1 1 9.5\n")
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=5) as b:
        t3 = b.autotune(1, cache=str(cache))
        assert t3["from_cache"] and b.it == 0
        assert (t3["packing"], t3["tile_rows"], t3["resident"], t3["chains_per_cu"]) == (other, 1024, 1, 1)
    _say("autotune cache: an entry the batch cannot run")
    with open(cache, "a") as fh:
        fh.write(f"{key}\t0 1536 1 1 1.0\n")  # 1536-row tiles need the running fit in LDS
    n_before = len(cache.read_text().splitlines())
    with sampler.BatchSampler(probs, m, nskip, nd, seeds, sweeps_per_launch=5) as b:
        t4 = b.autotune(1, cache=str(cache))
        assert not t4["from_cache"] and t4["sweeps"] == 13
    assert len(cache.read_text().splitlines()) == n_before + 1
    _say("autotune cache: another shape")
    with sampler.BatchSampler(_probs(ntree=8), 8, nskip, nd, seeds, sweeps_per_launch=5) as b:
        t5 = b.autotune(1, cache=str(cache))
        assert not t5["from_cache"] and t5["sweeps"] == 13
    last = cache.read_text().splitlines()[-1]
    assert "ntree=8" in last and last.split("\t")[0] != key
    with sampler.BatchSampler(probs, m, nskip, nd, seeds) as b:
        with pytest.raises(FbhError):
            b.autotune(1, cache=str(tmp_path / "no_such_dir" / "tune.txt"))
        assert b.it == 0


def test_gpu_launch_config_rejections():
    probs = _probs()[:2]
    with sampler.BatchSampler(probs, 8, 4, 2, [1, 2]) as b:
        with pytest.raises(FbhError):
            b.configure(3, 0)
        with pytest.raises(FbhError):
            b.configure(0, 512)
        with pytest.raises(FbhError):
            b.configure(0, 0, 3)
        with pytest.raises(FbhError):
            b.configure(0, 1536, 0, 1)  # 1536-row tiles with the fit streamed
        with pytest.raises(FbhError):
            b.autotune(1)  # 13 sweeps needed, 6 left
        assert b.it == 0
    big = [W.small_problem(12000, 2, 10, 7, "uniform", ntree=4)]
    with sampler.BatchSampler(big, 4, 2, 1, [1]) as b:
        assert (b.config["tile_rows"], b.config["resident"]) == (1024, 1)  # the running fit does not fit in LDS
        with pytest.raises(FbhError):
            b.configure(0, 1536)
        with pytest.raises(FbhError):
            b.configure(0, 1024, 0, 2)
        with pytest.raises(FbhError):
            b.autotune(1)  # one stage-1 candidate: 9 sweeps needed, 3 left
    with pytest.raises(FbhError):
        sampler.BatchSampler(big, 4, 2, 1, [1], packing=5)
    with pytest.raises(FbhError):
        sampler.BatchSampler(big, 4, 2, 1, [1], chains_per_cu=3)
    with pytest.raises(FbhError):
        sampler.BatchSampler(big, 4, 2, 1, [1], resident=2)
