"""Benchmark: tree-updates/sec/GPU of the batched plain sampler, plus min-ESS/sec.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batch512|fwbart_n4096|fpbart_n24576|fpbart_n24576_fx64|maxshape64|learned512|predict600|predict_probit100]

Default workload (the headline): batched fwbart -- 512 independent chains in one persistent launch, chain c with
n_c ~ U{3091..4573} rows, p_c ~ U{15..19} predictors, x ~ U(0,1) with 255 cutpoints, y = 10 sin(pi x1 x2) +
20 (x3-0.5)^2 + 10 x4 + 5 x5 + N(0, s_c^2), s_c ~ U(0.5, 2), w ~ U(0.5, 2); ntree 200, nskip 1000, ndpost 100.
A step is one sweep of every chain (m tree updates per chain plus the closing pass and the sigma / latent draw). W
warmup sweeps run untimed, then exactly K sweeps are timed; W + K covers the whole fit by default, so the kept draws of
the timed region give the min-ESS figure. Inputs are resident in HBM before the timed region starts. After the timed
region (nothing of it enters value) the headline line also carries the metric's second half, `learned_vs_plain`: the
learned512 comparison at 1,000 kept draws on the same chains (about a minute; --no-learned skips it).

Other workloads (extra lines, not the headline):
  fwbart_n4096   one chain, n = 4,096, p = 16, 100 cutpoints, ntree 200, nskip 1000, ndpost 100
  fpbart_n24576  one probit chain, n = 24,576, p = 3, ntree 50, nskip 1000, ndpost 100
  fpbart_n24576_fx64  the same chain with FX64 statistics and the automatic row split (8 workgroups), its latents
                 redrawn over all of them by the device-wide kernels (results: the oracle's FX64 twin)
  maxshape64     64 chains, n = 65,536, p = 32 (24 continuous + 8 integer columns), node cap 64, FX64 statistics,
                 ntree 200, nskip 500, ndpost 100 (default here: 5 warmup + 40 timed sweeps)
  learned512     the 512-chain batch twice: the plain fit, then the learned schedule -- plain sweeps, then sweeps with
                 BART's proposal in the lock-step pipeline collecting proposed births with the exact split conditional
                 as the training target, training of the network on the device, then sweeps with the mixture
                 0.2 uniform / 0.8 network up to the end (the kept draws); reports min-ESS/s of both (training time
                 apart), the network's held-out loss and the posterior-mean agreement
  predict600     predict_draw: 6 fits of the batched workload (200 trees, 100 kept draws each), 600 draws on 5,237
                 new rows; value in row-trees/s (a step is one pass over all 600 draws)
  predict_probit100  predict_draw for the 100 draws of the fpbart fit on its 24,576 rows (50 trees)

One JSON line on stdout. Statistics mode REF16 (double sums in 16 row lanes, bit-identical to the reference CPU
sampler) unless the workload says FX64. The CPU baseline is the reference's own core (oracle/_ref, built from the
reference sources) timed on this host, one chain per process, on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tree-updates/sec/GPU (batched chains) + min-ESS/sec, learned vs plain proposal"
HBM_PEAK_GBS = 8000.0

# name -> (chains, ntree, nskip, ndpost, probit, stats_mode, node_cap, default warmup, default timed steps or None)
WORKLOADS = {
    "batch512": (512, 200, 1000, 100, 0, 0, 0, 20, None),
    "fwbart_n4096": (1, 200, 1000, 100, 0, 0, 0, 20, None),
    "fpbart_n24576": (1, 50, 1000, 100, 1, 0, 0, 20, None),
    "fpbart_n24576_fx64": (1, 50, 1000, 100, 1, 1, 0, 20, None),
    "maxshape64": (64, 200, 500, 100, 0, 1, 64, 5, 40),
    "learned512": (512, 200, 1000, 100, 0, 0, 0, 20, None),
    "predict600": (6, 200, 1000, 100, 0, 0, 0, 3, 20),
    "predict_probit100": (1, 50, 1000, 100, 1, 0, 0, 3, 20),
    "text512": (512, 200, 1000, 100, 0, 0, 0, 0, None),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_chain(workload: str, c: int) -> dict:
    from fastbart_amd import workloads as Wk
    if workload in ("batch512", "learned512"):
        return Wk.batch512_chain(c)
    if workload == "fwbart_n4096":
        return Wk.workload_fwbart_n4096()
    if workload in ("fpbart_n24576", "fpbart_n24576_fx64"):
        return Wk.workload_fpbart_n24576()
    if workload == "maxshape64":
        return Wk.maxshape64_chain(c)
    raise ValueError(workload)


def describe(workload: str, chains: int, nskip: int, ndpost: int) -> str:
    return {
        "batch512": f"batched fwbart, {chains} chains, n_i~U{{3091..4573}}, p_i~U{{15..19}}, 255 cutpoints, ntree 200, "
                    f"nskip {nskip}, ndpost {ndpost}, REF16 statistics",
        "fwbart_n4096": f"fwbart, 1 chain, n 4096, p 16, 100 cutpoints, ntree 200, nskip {nskip}, ndpost {ndpost}, "
                        f"REF16 statistics",
        "fpbart_n24576": f"fpbart probit, 1 chain, n 24576, p 3, ntree 50, nskip {nskip}, ndpost {ndpost}, "
                         f"REF16 statistics",
        "fpbart_n24576_fx64": f"fpbart probit, 1 chain, n 24576, p 3, ntree 50, nskip {nskip}, ndpost {ndpost}, "
                              f"FX64 statistics, automatic row split",
        "maxshape64": f"max-shape batched fwbart, {chains} chains, n 65536, p 32, 255 cutpoints, node cap 64, ntree 200, "
                      f"nskip {nskip}, ndpost {ndpost}, FX64 statistics",
        "learned512": f"batched fwbart with the learned birth proposal, {chains} chains, n_i~U{{3091..4573}}, "
                      f"p_i~U{{15..19}}, 255 cutpoints, ntree 200, nskip {nskip}, ndpost {ndpost}, REF16 statistics",
    }[workload]


# ------------------------------------------------------------------ CPU baseline (runs before the GPU is touched)
def _cpu_worker(args):
    workload, chain, sweeps = args
    from oracle import pyoracle
    m, probit = WORKLOADS[workload][1], WORKLOADS[workload][4]
    pb = make_chain(workload, chain)
    core = pyoracle.RefCore() if pyoracle.RefCore.available() else pyoracle.Oracle()
    t0 = time.perf_counter()
    core.fit(pb, m, sweeps, 0, 1 + chain, keep_trees=0, probit=probit)
    return m * sweeps, time.perf_counter() - t0, os.getpid()


def host_cpu_info() -> dict:
    """The host the CPU baseline ran on: logical CPUs, the process's affinity, the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), cpu_model=model)


# The CPU baseline uses at most this many processes: the GPU host's CPU share for one GPU (the rest of the machine's
# cores serve other GPUs' jobs), whatever the machine's core count.
CPU_SHARE = 16


def baseline_cores() -> int:
    return min(CPU_SHARE, len(os.sched_getaffinity(0)))


def cpu_baseline(workload: str, cores: int):
    from oracle import pyoracle
    kind = "reference" if pyoracle.RefCore.available() else "port"
    if kind == "port" and not os.path.exists(pyoracle.ORACLE_SO):
        pyoracle.build(ref=False)
    nch = WORKLOADS[workload][0]
    if nch == 1:
        cores, jobs = 1, [(workload, 0, 400)]
    elif workload == "maxshape64":
        jobs = [(workload, c, 2) for c in range(cores)]
    else:
        jobs = [(workload, c, 40) for c in range(cores * 3)]
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    updates = sum(r[0] for r in res)
    cpu_s = sum(r[1] for r in res)
    # all-core throughput of the sampler itself: the busiest process's summed fit time (process start-up, imports and
    # data generation are outside the timed fits)
    per_proc = {}
    for r in res:
        per_proc[r[2]] = per_proc.get(r[2], 0.0) + r[1]
    makespan = max(per_proc.values())
    m = WORKLOADS[workload][1]
    note = " (REF16: the reference has no FX64 mode; same rows and trees)" if workload == "maxshape64" else ""
    return dict(value=updates / cpu_s if nch == 1 else updates / makespan, unit="tree-updates/s", cores=cores,
                kind=kind, host=host_cpu_info(), core_cap=f"at most {CPU_SHARE} processes: the host's CPU share "
                                                         f"for one GPU",
                sample=f"{len(jobs)} chain(s) of the workload x {jobs[0][2]} sweeps x {m} trees, one chain at a time "
                       f"per process on {cores} core(s) ({cpu_s:.1f} CPU-s of fits; busiest process {makespan:.2f} s; "
                       f"{wall:.1f} s wall with start-up; {updates / cpu_s:.0f} tree-updates/s per core){note}")


# ------------------------------------------------------------------ ESS (Geyer initial monotone sequence)
ESS_MIN_DRAWS = 100  # fewer kept draws leave Geyer's estimator at its floor for these strongly autocorrelated chains

def ess_geyer_cols(x: np.ndarray) -> np.ndarray:
    """ESS of every column of x [draws][series]: autocovariances by FFT, Geyer's initial monotone sequence of pair
    sums (Geyer 1992), ESS = n / (-1 + 2 sum of the monotone pair sums)."""
    x = np.asarray(x, np.float64)
    n, k = x.shape
    xc = x - x.mean(axis=0)
    var = (xc * xc).sum(axis=0) / n
    f = np.fft.rfft(xc, 2 * n, axis=0)
    acov = np.fft.irfft(f * np.conj(f), axis=0)[:n] / n
    ok = var > 0
    rho = np.where(ok, acov / np.where(ok, var, 1.0), 0.0)
    m = (n - 1) // 2
    if m < 1:  # fewer than three draws: no pair sums
        return np.full(k, float(n))
    pairs = rho[0:2 * m:2] + rho[1:2 * m:2]               # [m][k]
    pos = pairs > 0
    first_neg = np.where(pos.all(axis=0), m, np.argmin(pos, axis=0))
    mono = np.minimum.accumulate(pairs, axis=0)
    keep = np.arange(m)[:, None] < first_neg[None, :]
    tau = -1.0 + 2.0 * (mono * keep).sum(axis=0)
    ess = n / np.maximum(tau, 1e-12)
    return np.where(ok, ess, float(n))


def ess_geyer(x: np.ndarray) -> float:
    return float(ess_geyer_cols(np.asarray(x, np.float64)[:, None])[0])


def min_ess_chain(sigma_kept, yhat: np.ndarray, max_rows: int = 0) -> float:
    """min over sigma (wbart) and every f(x_i) column of the chain's kept draws."""
    cols = yhat if not max_rows else yhat[:, np.linspace(0, yhat.shape[1] - 1, max_rows).astype(int)]
    e = float(ess_geyer_cols(cols).min())
    return e if sigma_kept is None else min(ess_geyer(sigma_kept), e)


# ------------------------------------------------------------------ measured HBM traffic
# sweeps the autotuner takes at one sweep per trial launch: a warm-up launch, two rounds of at most 2 builds x 3 tile
# kinds, two rounds of 3 packings
TUNE_SWEEPS_MAX = 1 + 2 * 6 + 2 * 3


TRAFFIC_SOURCE = ("profiles/k_sweeps_{}_traffic.txt: FETCH_SIZE + WRITE_SIZE per sweep from two rocprofv3 --pmc "
                  "runs of this workload, calibrated against a plain copy kernel, scaled to this run's sweeps per "
                  "launch (not measured in this run)")


def measured_traffic(workload: str, sweeps_per_launch: float):
    """HBM bytes per launch of the sweep kernel from the committed counter summary (FETCH_SIZE + WRITE_SIZE per sweep,
    each scaled by the ratio the same counter reads for a plain copy of known size), or None."""
    t = {}
    try:
        with open(os.path.join(ROOT, "profiles", f"k_sweeps_{workload}_traffic.txt")) as f:
            for ln in f:
                k, sep, v = ln.partition(":")
                if sep and k.strip() in ("fetch_bytes_per_sweep", "write_bytes_per_sweep"):
                    t[k.strip()] = float(v.split()[0])
        return (t["fetch_bytes_per_sweep"] + t["write_bytes_per_sweep"]) * sweeps_per_launch
    except (OSError, KeyError, ValueError, IndexError):
        return None


# ------------------------------------------------------------------ learned vs plain
def learned_vs_plain(args, chains, nskip, ndpost, cpu):
    """Plain fit and learned schedule on the same chains and seeds; prints the learned512 line."""
    print(json.dumps(learned_compare(args, chains, nskip, ndpost, cpu)), flush=True)
    return 0


def learned_summary(line):
    """The learned-vs-plain figures the headline line carries (the learned512 line has the rest)."""
    L, P = line["learned"], line["plain"]
    keys = ("ess_sigma_median", "ess_f_p5_median", "ess_f_median_median", "ess_f_min_median", "sum_min_ess")
    return {
        "what": (f"the learned birth proposal against the plain GPU sampler on the same {line['config']['chains']} "
                 f"chains and seeds, {line['ndpost']} kept draws each (bench.py --workload learned512 "
                 f"--ndpost {line['ndpost']} prints the whole line)"),
        "schedule": line["config"]["schedule"],
        "plain_min_ess_per_sec": P["min_ess_per_sec"], "learned_min_ess_per_sec": L["min_ess_per_sec"],
        "ratio": line["ess_per_sec_ratio_learned_over_plain"], "min_ess_note": line["min_ess_note"],
        "posterior_mean_corr": line["posterior_mean_corr"],
        "seconds": {"plain": P["seconds"], "learned_sampling": L["seconds_sampling"],
                    "learned_training": L["seconds_training"]},
        "ms_per_learned_sweep": L["ms_per_learned_sweep"],
        "heldout_loss_nats": {"network": L["heldout_loss"], "untrained": L["heldout_loss_untrained"],
                              "floor_entropy": L["heldout_entropy_floor"], "uniform_rule": L["heldout_uniform_rule"]},
        "accepted_births": {"plain": line["plain_births"], "learned": L["births"]},
        "sigma_mean_kept": line["sigma_mean_kept"],
        "ess_per_chain_median": {"plain": {k: P[k] for k in keys}, "learned": {k: L[k] for k in keys}},
    }


def learned_compare(args, chains, nskip, ndpost, cpu):
    """Plain fit and learned schedule on the same chains and seeds; min-ESS/s of both over the kept draws."""
    from fastbart_amd import sampler
    from fastbart_amd.mlp import MLP
    total = nskip + ndpost
    n_plain0 = max(0, nskip - args.collect - args.learned_burn)
    probs = [make_chain("learned512", c) for c in range(chains)]
    seeds = [1 + c for c in range(chains)]
    rows_total = sum(p["q"].shape[1] for p in probs)
    M = 200

    def ess_of(b):
        """(sum over chains of the min-ESS, posterior means, ESS distribution: per chain ESS(sigma) and the 5th / 50th
        percentile and minimum of ESS(f(x_i)) over its rows)"""
        tot = 0.0
        means, es, f5, f50, fmin = [], [], [], [], []
        for c in range(chains):
            out = b.fetch(c)
            yh = b.yhat(c)
            cols = yh if not args.ess_rows else yh[:, np.linspace(0, yh.shape[1] - 1, args.ess_rows).astype(int)]
            ef = ess_geyer_cols(cols)
            esig = ess_geyer(out["sigma"][nskip:])
            tot += min(esig, float(ef.min()))
            means.append(out["trmean"].astype(np.float64))
            es.append(esig)
            f5.append(float(np.percentile(ef, 5)))
            f50.append(float(np.median(ef)))
            fmin.append(float(ef.min()))
        dist = {"ess_sigma_median": float(np.median(es)), "ess_f_p5_median": float(np.median(f5)),
                "ess_f_median_median": float(np.median(f50)), "ess_f_min_median": float(np.median(fmin))}
        return tot, means, dist

    # plain (the launch configuration the autotuner picks for this batch: serpentine packing, the automatic build and
    # tile kind)
    b = sampler.BatchSampler(probs, M, nskip, ndpost, seeds, keep_trees=1, sweeps_per_launch=20, packing=1, graph=1)
    b.sync()
    t0 = time.perf_counter()
    b.run(total)
    b.sync()
    t_plain = time.perf_counter() - t0
    ess_plain, mean_plain, dist_plain = ess_of(b)
    births_plain = int(sum(int(b.fetch(c)["stats"][0]) for c in range(chains)))
    sigma_plain = float(np.mean([b.fetch(c)["sigma"][nskip:].astype(np.float64).mean() for c in range(chains)]))
    b.close()
    log(f"plain: {total} sweeps {t_plain:.2f} s, sum min-ESS {ess_plain:.1f}")
    # learned schedule
    net = MLP((513, 256, 256, 8160), max_batch=max(chains, 512), seed=5)
    b = sampler.BatchSampler(probs, M, nskip, ndpost, seeds, keep_trees=1, sweeps_per_launch=20, packing=1, graph=1)
    b.sync()
    t0 = time.perf_counter()
    b.run(n_plain0)
    b.learned_setup(net, 0.0, pair_capacity=args.pairs)
    cond = args.target == "conditional"
    b.learned_target(1 if cond else 0, fraction=args.pair_fraction if cond else 1.0, seed=11)
    b.run_learned(args.collect)
    b.sync()
    t_a = time.perf_counter() - t0
    npairs = b.learned_pairs()
    ent, lall = b.learned_target_stats()
    # the last tenth of the pairs (the latest sweeps) stays out of training: the loss there is out of sample
    hold = npairs // 10
    b.learned_holdout(hold)
    held0 = b.learned_eval(npairs - hold, hold)[0] if hold else None
    t1 = time.perf_counter()
    losses = b.learned_train(args.train_steps, batch=512, lr=args.lr, seed=3)
    t_train = time.perf_counter() - t1
    held = b.learned_eval(npairs - hold, hold) if hold else (None, None, None)
    b.learned_setup(net, args.alpha, pair_capacity=0)
    t2 = time.perf_counter()
    b.run_learned(total - n_plain0 - args.collect)
    b.sync()
    t_b = time.perf_counter() - t2
    t_learned = t_a + t_b
    ms_sweep, nls = b.learned_timing()
    ess_learned, mean_learned, dist_learned = ess_of(b)
    st = np.array([b.fetch(c)["stats"] for c in range(chains)])
    sigma_learned = float(np.mean([b.fetch(c)["sigma"][nskip:].astype(np.float64).mean() for c in range(chains)]))
    b.close()
    fwd_ms, _ = net.timing() if hasattr(net, "timing") else (0.0, 0)
    corr = float(np.mean([np.corrcoef(a, bb)[0, 1] for a, bb in zip(mean_plain, mean_learned)]))
    loss0, loss1 = float(losses[:20].mean()), float(losses[-20:].mean())
    log(f"learned: {t_learned:.2f} s sampling + {t_train:.2f} s training on {npairs - hold} pairs "
        f"(training loss {loss0:.3f} -> {loss1:.3f}; held-out loss {held0} -> {held[0]}, "
        f"its floor {held[1]} and uniform level {held[2]} nats); "
        f"sum min-ESS {ess_learned:.1f}; "
        f"posterior-mean correlation with plain {corr:.4f}")
    value = chains * M * total / t_learned
    ess_ps, ess_pp = ess_learned / t_learned, ess_plain / t_plain
    ess_note = None
    if ndpost < ESS_MIN_DRAWS:
        ess_ps = ess_pp = None
        ess_note = (f"not computed: {ndpost} kept draws per chain; Geyer's estimator needs at least {ESS_MIN_DRAWS} "
                    f"(bench.py --workload learned512 --ndpost 1000)")
    line = {
        "metric": METRIC, "value": value, "unit": "tree-updates/s", "n_gpus": 1, "steps": total, "warmup": 0,
        "ms_per_step": t_learned / total * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": describe("learned512", chains, nskip, ndpost), "chains": chains, "ntree": M,
                   "rows_total": rows_total, "parallelism": "chains",
                   "schedule": f"{n_plain0} plain sweeps, {args.collect} sweeps collecting births (alpha 0), "
                               f"{args.train_steps} Adam steps, {total - n_plain0 - args.collect} sweeps at alpha "
                               f"{args.alpha}"},
        "min_ess_per_sec": ess_ps,
        "min_ess_note": ess_note,
        "ndpost": ndpost,
        "plain": {"min_ess_per_sec": ess_pp, "seconds": t_plain, "sum_min_ess": ess_plain, **dist_plain},
        "learned": {"min_ess_per_sec": ess_ps, "seconds_sampling": t_learned, "seconds_training": t_train,
                    "sum_min_ess": ess_learned, **dist_learned, "pairs": npairs, "ms_per_learned_sweep": ms_sweep,
                    "learned_sweeps": nls, "births": int(st[:, 0].sum()), "deaths": int(st[:, 1].sum()),
                    "target": args.target, "loss_first20": loss0, "loss_last20": loss1,
                    "heldout_pairs": hold, "heldout_loss_untrained": held0, "heldout_loss": held[0],
                    "heldout_entropy_floor": held[1] if held[1] is not None and np.isfinite(held[1]) else None,
                    "heldout_uniform_rule": held[2] if held[2] is not None and np.isfinite(held[2]) else None,
                    "target_entropy_nats": ent if np.isfinite(ent) else None,
                    "uniform_rule_nats": lall if np.isfinite(lall) else None},
        "plain_births": births_plain,
        # the chains' sigma still falls for tens of thousands of sweeps (burn-in): over the kept draws a lower mean
        # is a chain further converged
        "sigma_mean_kept": {"plain": sigma_plain, "learned": sigma_learned},
        "ess_per_sec_ratio_learned_over_plain": ess_ps / ess_pp if ess_ps is not None and ess_pp else None,
        "posterior_mean_corr": corr,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    return line


# ------------------------------------------------------------------ prediction (compile_trees / predict_draw)
def _predict_cpu(jobs, seconds):
    """The oracle's restatement of predict_draw_cpp (src/predict.cpp:49-73) on one core: single-draw calls spread
    over the fits for about `seconds`; returns (row-trees per second, CPU seconds, calls)."""
    from oracle import pyoracle
    orc = pyoracle.Oracle()
    prepared = []
    for text, cuts, x, mu in jobs:
        oc = orc.compile_trees(text, cuts)
        prepared.append((oc, [np.ascontiguousarray(x[:, v]) for v in range(x.shape[1])], mu))
    work, k, t0 = 0, 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        oc, cols, mu = prepared[k % len(prepared)]
        orc.predict_draw(oc, cols, 1 + (k // len(prepared)) % oc["ndpost"], mu)
        work += len(cols[0]) * oc["ntree"]
        k += 1
    dt = time.perf_counter() - t0
    return work / dt, dt, k


def predict_bench(args, wl):
    """Fits of the workload's chains (full nskip + ndpost on the device), their treedraws text compiled once, then
    predict_draw for every kept draw on new rows resident on the device. value = rows x trees per second of the
    batched path (one launch per fit over its draws, device-side output); the line also carries single-draw calls
    (one launch per draw) and the host-output rate (PCIe included)."""
    from fastbart_amd import sampler
    from fastbart_amd.predict import CompiledTrees, DeviceBuffer, DeviceRows, device_sync
    from fastbart_amd import workloads as Wk
    if wl == "predict600":
        fits = [Wk.batch512_chain(c) for c in range(6)]
        m, nskip, ndpost, probit, n_new = 200, 1000, 100, 0, 5237
        desc = ("predict_draw on 6 fits of the batched fwbart workload (chains 0-5, ntree 200, 100 kept draws each): "
                "600 draws x 5237 new rows, x ~ U(0,1)")
    else:
        fits = [Wk.workload_fpbart_n24576()]
        m, nskip, ndpost, probit, n_new = 50, 1000, 100, 1, 24576
        desc = "predict_draw for all 100 draws of the fpbart probit fit (ntree 50) on its 24576 training rows"
    t0 = time.perf_counter()
    with sampler.BatchSampler(fits, m, nskip, ndpost, [1 + c for c in range(len(fits))], probit=probit,
                              keep_trees=1) as b:
        b.run(nskip + ndpost)
        outs = [b.fetch(c) for c in range(len(fits))]
    log(f"{wl}: fitted {len(fits)} chain(s) in {time.perf_counter() - t0:.1f} s")
    cts, rows, hostx, jobs = [], [], [], []
    for c, pb in enumerate(fits):
        p = pb["q"].shape[0]
        text, _ = sampler.tree_text(outs[c]["trees"], p)
        mu = float(pb["offset"]) if probit else float(pb["fmean"])
        ct = CompiledTrees(text, pb["cutpoints"], mu=mu)
        if wl == "predict600":
            x = np.random.Generator(np.random.PCG64(4000 + c)).uniform(0.0, 1.0, size=(n_new, p))
        else:
            x = pb["x"]
        cts.append(ct)
        rows.append(DeviceRows(x))
        hostx.append(x)
        jobs.append((text, pb["cutpoints"], x, mu))
    nodes = [ct.nnodes / ct.ndpost for ct in cts]
    rank = all(ct.rank_form for ct in cts)
    d_draws = DeviceBuffer.from_array(np.arange(1, ndpost + 1, dtype=np.int32))
    d_outs = [DeviceBuffer(8 * ndpost * r.n) for r in rows]
    draws_total = ndpost * len(fits)
    rowtrees = sum(ndpost * r.n * m for r in rows)

    def batched():
        for ct, r, o in zip(cts, rows, d_outs):
            r.predict_device(ct, d_draws, ndpost, o)
        device_sync()

    singles = [DeviceBuffer.from_array(np.array([k], np.int32)) for k in range(1, ndpost + 1)]

    def single_calls():
        for ct, r, o in zip(cts, rows, d_outs):
            for k in range(ndpost):
                r.predict_device(ct, singles[k], 1, o)
        device_sync()

    for _ in range(max(1, args.warmup if args.warmup is not None else 3)):
        batched()
        single_calls()
    reps = max(1, args.steps if args.steps is not None else 20)
    k0 = [ct.timing() for ct in cts]
    t1 = time.perf_counter()
    for _ in range(reps):
        batched()
    T = time.perf_counter() - t1
    k1 = [ct.timing() for ct in cts]
    # average kernel duration of the timed batched launches (HIP events on the launch stream)
    ms_sum = sum(a1 * n1 - a0 * n0 for (a0, n0), (a1, n1) in zip(k0, k1))
    nl = sum(n1 - n0 for (a0, n0), (a1, n1) in zip(k0, k1))
    ms_launch = ms_sum / nl if nl else 0.0
    t2 = time.perf_counter()
    for _ in range(reps):
        single_calls()
    T1 = time.perf_counter() - t2
    t3 = time.perf_counter()
    for _ in range(max(1, reps // 4)):
        for ct, r in zip(cts, rows):
            r.predict(ct, np.arange(1, ndpost + 1))
    T2 = (time.perf_counter() - t3) / max(1, reps // 4)
    value = rowtrees * reps / T
    # algorithmic bytes per launch (one fit, ndpost draws): x once and f once per (draw, row), 24 B per node of the
    # draw's trees
    p_avg = float(np.mean([r.p for r in rows]))
    n_avg = float(np.mean([r.n for r in rows]))
    bytes_launch = ndpost * n_avg * (8 * p_avg + 8) + ndpost * 24 * float(np.mean(nodes))
    achieved = bytes_launch / (ms_launch / 1e3) / 1e9 if ms_launch else 0.0
    log(f"{wl}: batched {value / 1e9:.2f} G row-trees/s ({T / reps * 1e3:.3f} ms per pass of {draws_total} draws, "
        f"kernel {ms_launch:.3f} ms per launch of {ndpost} draws); single-draw calls "
        f"{rowtrees * reps / T1 / 1e9:.2f} G/s; host output {rowtrees / T2 / 1e9:.2f} G/s; rank form {rank}")
    cpu = None
    if not args.no_cpu:
        cpu_v, cpu_s, calls = _predict_cpu(jobs, 5.0)
        cpu = dict(value=cpu_v, unit="row-trees/s", cores=1, kind="port",
                   sample=f"{calls} single-draw predict_draw calls of the oracle's restatement of src/predict.cpp "
                          f"(same fits and rows), one core, {cpu_s:.1f} s")
    for x in d_outs + singles + [d_draws]:
        x.close()
    if args.calib_copy:
        ms = sampler.debug_copy(args.calib_copy)
        log(f"calibration copy of {args.calib_copy} MiB: {ms:.3f} ms")
    line = {
        "metric": "predict_draw row-trees/sec (compile_trees / predict_draw)", "value": value, "unit": "row-trees/s",
        "n_gpus": 1, "steps": reps, "warmup": args.warmup if args.warmup is not None else 3,
        "ms_per_step": T / reps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": desc, "fits": len(fits), "draws": draws_total, "rows": [r.n for r in rows],
                   "ntree": m, "nodes_per_draw": float(np.mean(nodes)), "kernel": "k_predict_rank" if rank else
                   "k_predict", "parallelism": "rows x draws"},
        "single_draw_calls": {"value": rowtrees * reps / T1, "ms_per_call": T1 / (reps * draws_total) * 1e3},
        "host_output": {"value": rowtrees / T2, "note": "fbh_predict_rows: device result copied to host memory"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_predict_rank" if rank else "k_predict", "kernel_avg_ms": ms_launch, "launches": nl,
                     "algorithmic_bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)
    for r in rows:
        r.close()
    for ct in cts:
        ct.close()
    return 0


# ------------------------------------------------------------------ treedraws text of a batched fit
def text_bench(args):
    """The output half of fbart_cpp for every chain of the 512-chain workload: the fit runs on the device (not
    timed), each chain's kept trees are copied to the host, then formatted as treedraws text with varcount by the
    library's threaded writer (src/glue.cpp:32-51). value = trees formatted per second over all chains (10.24 M
    trees); the copies to the host are timed apart."""
    from fastbart_amd import sampler
    from fastbart_amd import workloads as Wk
    nch, M, NSKIP, NDPOST = 512, 200, 1000, 100
    chains = min(nch, args.chains) if args.chains else nch
    probs = [Wk.batch512_chain(c) for c in range(chains)]
    t0 = time.perf_counter()
    with sampler.BatchSampler(probs, M, NSKIP, NDPOST, [1 + c for c in range(chains)], keep_trees=1,
                              chains_per_cu=2 if chains > 256 else 0, packing=1) as b:
        b.run(NSKIP + NDPOST)
        b.sync()
        log(f"text512: fitted {chains} chains in {time.perf_counter() - t0:.1f} s")
        t_fetch = t_fmt = 0.0
        nbytes = 0
        sample = []
        for c in range(chains):
            t1 = time.perf_counter()
            trees = b.fetch(c)["trees"]
            t2 = time.perf_counter()
            text, vc = sampler.tree_text(trees, probs[c]["q"].shape[0])
            t3 = time.perf_counter()
            t_fetch += t2 - t1
            t_fmt += t3 - t2
            nbytes += len(text)
            if c < 4:
                sample.append((trees, probs[c]["q"].shape[0], text))
    ntrees = chains * NDPOST * M
    value = ntrees / t_fmt
    log(f"text512: {ntrees / 1e6:.2f} M trees, {nbytes / 1e6:.0f} MB of text in {t_fmt:.2f} s "
        f"({value / 1e6:.2f} M trees/s); host copies of the kept trees {t_fetch:.2f} s")
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle
        o = pyoracle.Oracle()
        t4 = time.perf_counter()
        k = 0
        while time.perf_counter() - t4 < 5.0 or k == 0:
            trees, p, text = sample[k % len(sample)]
            ref, _ = o.tree_text(trees, p)
            assert ref == text, "treedraws text differs from the oracle's"
            k += 1
        cs = time.perf_counter() - t4
        cpu = dict(value=k * NDPOST * M / cs, unit="trees/s", cores=1, kind="port",
                   sample=f"{k} chains' treedraws text (20,000 trees each) by the oracle's restatement of "
                          f"src/glue.cpp:32-51 (snprintf %.10g), one core, {cs:.1f} s; the texts equal the library's")
    line = {
        "metric": "treedraws text trees/sec (output half of fbart_cpp, all chains of the batched fit)",
        "value": value, "unit": "trees/s", "n_gpus": 1, "steps": chains, "warmup": 0,
        "ms_per_step": t_fmt / chains * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 leaf values as %.10g", "data": "synthetic",
        "config": {"workload": describe("batch512", chains, NSKIP, NDPOST) + "; treedraws text of every kept draw",
                   "chains": chains, "trees": ntrees, "text_bytes": nbytes, "host_threads": os.cpu_count()},
        "host_copy_s": t_fetch,
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


# ------------------------------------------------------------------ main
def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed sweeps (default: the rest of the fit)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="batch512", choices=sorted(WORKLOADS))
    ap.add_argument("--chains", type=int, default=None, help="chains of a batched workload (default: all)")
    ap.add_argument("--sweeps-per-launch", type=int, default=20,
                    help="sweeps per persistent launch (the warmup and the timed sweeps split into equal launches)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="batch512: skip the learned-vs-plain comparison the headline line carries")
    ap.add_argument("--learned-ndpost", type=int, default=1000,
                    help="batch512: kept draws of the learned-vs-plain comparison")
    ap.add_argument("--ess-rows", type=int, default=0, help="rows per chain used for ESS (0 = all)")
    ap.add_argument("--row-split", type=int, default=None,
                    help="workgroups per chain for FX64 workloads (-1 = one workgroup per CU; default: -1 for "
                         "maxshape64, else 1)")
    ap.add_argument("--packing", default="auto",
                    help="workgroup -> chain packing of the sweep kernel: 0 batch order, 1 serpentine, 2 largest "
                         "first, auto = the autotuner picks packing, tile rows and build during the warmup")
    ap.add_argument("--tune-cache", default=None,
                    help="autotuner cache file: a batch shape tuned before takes the stored choice without trials")
    ap.add_argument("--tile-rows", type=int, default=0, help="rows per LDS tile (0 automatic, 1024, 1536)")
    ap.add_argument("--resident", type=int, default=0,
                    help="running fit during a launch: 1 streamed, 2 in LDS, 0 automatic")
    ap.add_argument("--chains-per-cu", type=int, default=0,
                    help="sweep-kernel build: 1 (168 VGPRs) or 2 (128 VGPRs, two workgroups per CU); 0 automatic")
    ap.add_argument("--no-graph", action="store_true", help="plain stream launches instead of a captured hipGraph")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration under rocprofv3)")
    ap.add_argument("--ndpost", type=int, default=None, help="kept sweeps (learned512; longer runs steady ESS)")
    ap.add_argument("--collect", type=int, default=100, help="learned512: sweeps collecting births")
    ap.add_argument("--learned-burn", type=int, default=100, help="learned512: burn-in sweeps at the mixture")
    ap.add_argument("--train-steps", type=int, default=4000)
    ap.add_argument("--pairs", type=int, default=600000, help="learned512: training-pair capacity")
    ap.add_argument("--pair-fraction", type=float, default=0.12,
                    help="learned512, conditional target: share of the proposed births taken as pairs")
    ap.add_argument("--alpha", type=float, default=0.8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--target", choices=("conditional", "accepted"), default="conditional",
                    help="learned512 training pairs: every proposed birth's node labelled from the exact split "
                         "conditional, or each accepted birth with BART's rule")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every chain runs on a single MI355X (chains would shard across GPUs "
              "with no exchange between them). Use --gpus 1.", file=sys.stderr)
        return 2
    wl = args.workload
    if wl.startswith("predict"):
        return predict_bench(args, wl)
    if wl == "text512":
        return text_bench(args)
    nch, M, NSKIP, NDPOST, probit, stats_mode, node_cap, w_def, k_def = WORKLOADS[wl]
    chains = min(nch, args.chains) if args.chains else nch
    if wl == "learned512":
        cpu = None
        if not args.no_cpu:
            cpu = cpu_baseline("batch512", baseline_cores())
        return learned_vs_plain(args, chains, NSKIP, args.ndpost or NDPOST, cpu)
    W_ = max(0, args.warmup if args.warmup is not None else w_def)
    if args.packing == "auto":
        # the autotuner's trials are warmup sweeps: the warmup is at least as long as the longest tuning
        W_ = max(W_, TUNE_SWEEPS_MAX)
    K = args.steps if args.steps is not None else (k_def if k_def is not None else NSKIP + NDPOST - W_)
    total = W_ + K
    nskip = max(0, total - NDPOST) if total >= NDPOST else 0
    ndpost = total - nskip
    spl = max(1, min(args.sweeps_per_launch, max(W_, 1), K))

    cpu = None
    if not args.no_cpu:
        cores = baseline_cores()
        log(f"cpu baseline on up to {cores} cores ...")
        cpu = cpu_baseline(wl, cores)
        log("cpu baseline:", cpu)

    from fastbart_amd import sampler
    from fastbart_amd._lib import FbhError

    t0 = time.perf_counter()
    probs = [make_chain(wl, c) for c in range(chains)]
    seeds = [1 + c for c in range(chains)]
    rows_total = sum(p["q"].shape[1] for p in probs)
    log(f"{wl}: generated {chains} chains in {time.perf_counter() - t0:.1f} s; total rows {rows_total}")
    split = args.row_split if args.row_split is not None else (-1 if wl in ("maxshape64", "fpbart_n24576_fx64") else 1)
    tune = args.packing == "auto"
    b = sampler.BatchSampler(probs, M, nskip, ndpost, seeds, probit=probit, keep_trees=1, stats_mode=stats_mode,
                             node_cap=node_cap, sweeps_per_launch=spl, row_split=split,
                             packing=0 if tune else int(args.packing), tile_rows=args.tile_rows,
                             graph=0 if args.no_graph else 1, chains_per_cu=args.chains_per_cu,
                             resident=args.resident)
    split = b.row_split
    b.sync()
    t_w = time.perf_counter()
    tuned = None
    if tune and W_:
        # the autotuner's trials are real warmup sweeps (one sweep per trial launch); what is left of the warmup runs
        # in the chosen configuration
        try:
            tuned = b.autotune(1, cache=args.tune_cache)
        except FbhError as e:
            log("autotune skipped:", e)
        if tuned:
            log("autotune:", {k: tuned[k] for k in ("packing", "tile_rows", "resident", "chains_per_cu", "sweeps",
                                                    "from_cache")},
                [(c["packing"], c["tile_rows"], c["resident"], c["chains_per_cu"], round(c["ms_per_sweep"], 3))
                 for c in tuned["candidates"]])
    if W_ > b.it:
        b.run(W_ - b.it)
    b.sync()
    t_w = time.perf_counter() - t_w
    log(f"warmup {W_} sweeps: {t_w:.2f} s")
    launch_cfg = b.config
    redraw = b.probit_info["mode"] if probit else 0
    avg0, nl0, _ = b.timing()
    t1 = time.perf_counter()
    b.run(K)
    b.sync()
    T = time.perf_counter() - t1
    if b.it != W_ + K:
        raise RuntimeError(f"ran {b.it} sweeps, expected {W_} warmup + {K} timed")
    avg_ms, nl, _ = b.timing()
    value = chains * M * K / T
    log(f"timed {K} sweeps: {T:.3f} s -> {value / 1e6:.3f} M tree-updates/s; kernel avg {avg_ms:.3f} ms over {nl} "
        f"launches")

    # roofline of the sweep kernel over the timed launches: algorithmic bytes (37 per row per tree update, times
    # (m+1)/m for the closing pass) per launch / average launch duration (HIP events on the sampler's stream)
    bytes_per_sweep = 37.0 * rows_total * (M + 1)
    nlt = nl - nl0
    ms_launch = (avg_ms * nl - avg0 * nl0) / nlt if nlt else 0.0
    sweeps_launch = K / nlt if nlt else 0.0
    achieved = bytes_per_sweep * sweeps_launch / (ms_launch / 1e3) / 1e9 if nlt else 0.0
    traffic = measured_traffic(wl, sweeps_launch)

    ess, ess_note = None, None
    if ndpost < ESS_MIN_DRAWS:
        ess_note = (f"not computed: {ndpost} kept draws per chain; Geyer's estimator needs at least {ESS_MIN_DRAWS} "
                    f"for a figure that is not at its floor (run the whole fit: no --steps)")
    else:
        t2 = time.perf_counter()
        tot = 0.0
        for c in range(chains):
            out = b.fetch(c)
            yh = b.yhat(c)
            tot += min_ess_chain(None if probit else out["sigma"][nskip:], yh, args.ess_rows)
        ess = tot / (t_w + T)
        ess_note = (f"sum over chains of min(ESS(sigma), min_i ESS(f(x_i))) over {ndpost} kept draws, Geyer's initial "
                    f"monotone sequence estimator, / (warmup + timed) wall time")
        log(f"min-ESS: sum over chains {tot:.1f} ({time.perf_counter() - t2:.1f} s to compute)")
    b.close()
    if args.calib_copy:
        # a plain copy kernel that reads and writes calib_copy MiB once each
        ms = sampler.debug_copy(args.calib_copy)
        log(f"calibration copy of {args.calib_copy} MiB: {ms:.3f} ms, {2 * (args.calib_copy << 20) / ms / 1e6:.0f} GB/s")

    line = {
        "metric": METRIC, "value": value, "unit": "tree-updates/s", "n_gpus": 1, "steps": K, "warmup": W_,
        "ms_per_step": T / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": describe(wl, chains, nskip, ndpost), "chains": chains, "ntree": M,
                   "rows_total": rows_total,
                   "parallelism": "chains" if split <= 1 else f"chains x {split} workgroups (row split)",
                   "packing": launch_cfg["packing"], "tile_rows": launch_cfg["tile_rows"],
                   "fit_in_lds": launch_cfg["resident"] == 2,
                   "chains_per_cu": launch_cfg["chains_per_cu"], "graph": bool(launch_cfg["graph"]),
                   "autotuned": tuned is not None,
                   "tune_from_cache": bool(tuned and tuned["from_cache"]),
                   "probit_redraw": ({1: "in the sweep kernel", 2: "spread over the device"}[redraw]
                                     if probit else None)},
        "warmup_requested": args.warmup,
        "warmup_note": (f"{W_} warmup sweeps run: the autotuner's trials take up to {TUNE_SWEEPS_MAX} of them"
                        if W_ != (args.warmup if args.warmup is not None else w_def) else None),
        "min_ess_per_sec": ess,
        "min_ess_note": ess_note,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": TRAFFIC_SOURCE.format(wl) if traffic is not None else None,
                     "kernel": "k_sweeps" if redraw != 2 else "k_sweeps + k_pr_* (the latent redraw)",
                     "kernel_avg_ms": ms_launch, "launches": nlt,
                     "sweeps_per_launch": sweeps_launch, "bytes_per_tree_update_row": 37,
                     "algorithmic_bytes_per_launch": bytes_per_sweep * sweeps_launch},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["gpu_over_cpu"] = value / cpu["value"]
    if wl == "batch512" and not args.no_learned:
        # the metric's second half: min-ESS/s of the learned proposal against the plain sampler, on the same chains
        # (after the timed region; nothing of it enters value)
        try:
            rows_keep = args.ess_rows
            args.ess_rows = args.ess_rows or 400
            lv = learned_compare(args, chains, NSKIP, args.learned_ndpost, None)
            args.ess_rows = rows_keep
            line["learned_vs_plain"] = learned_summary(lv)
        except Exception as e:  # the headline line is printed whatever happens here
            line["learned_vs_plain"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
