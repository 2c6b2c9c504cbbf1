"""Proposal study on the CPU (not collected by pytest): how much can any learned birth proposal gain on the batched
fwbart workload's chains?

A learned birth proposal maps a node's features to a distribution over rules (v, c); the best it can approach is the
rule's exact conditional given the tree, sigma and the partial residuals. The oracle's study mode draws births from
that conditional with probability alpha (else BART's uniform rule), the mixture's probability entering the
Metropolis-Hastings ratios (exactness by enumeration: tests/test_tree_posterior.py). This script runs chains of the
workload (n_i ~ U{3091..4573}, p_i ~ U{15..19}, ntree 200) with and without it, on the same data and seeds, and
reports the per-chain ESS of sigma and of f(x_i) over the kept draws (Geyer's initial monotone sequence estimator, as
bench.py) with bootstrap intervals over chains.

    python tests/study_informed_rule.py --chains 16 --nskip 1000 --ndpost 1000 --alpha 0.8
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for d in (ROOT, os.path.join(ROOT, "oracle")):
    if d not in sys.path:
        sys.path.insert(0, d)

M = 200


def one_chain(job):
    c, alpha, nskip, ndpost = job
    import bench
    from fastbart_amd import workloads as W
    from pyoracle import Oracle
    o = Oracle()
    pb = W.batch512_chain(c)
    ch = o.chain(pb, M, nskip, ndpost, 1 + c)
    if alpha > 0:
        ch.set_informed(alpha)
    t = time.perf_counter()
    ch.sweeps(nskip + ndpost)
    sec = time.perf_counter() - t
    births, deaths = (int(x) for x in ch.state()["stats"][:2])
    yhat = o.yhat(ch.trees_out.reshape(ndpost, M, -1), pb["q"])
    ef = bench.ess_geyer_cols(yhat)
    es = bench.ess_geyer(ch.sigma_out[nskip:].astype(np.float64))
    out = dict(chain=c, alpha=alpha, sec=sec, births=births, deaths=deaths, informed_births=ch.informed_births,
               ess_sigma=es, ess_f_p5=float(np.percentile(ef, 5)), ess_f_median=float(np.median(ef)),
               ess_f_min=float(ef.min()), min_ess=min(es, float(ef.min())), post_mean=yhat.mean(axis=0).tolist())
    ch.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--chains", type=int, default=16)
    ap.add_argument("--nskip", type=int, default=1000)
    ap.add_argument("--ndpost", type=int, default=1000)
    ap.add_argument("--alpha", type=float, default=0.8)
    ap.add_argument("--workers", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--out", default=None, help="per-chain results as JSON")
    a = ap.parse_args()
    jobs = [(c, al, a.nskip, a.ndpost) for c in range(a.chains) for al in (0.0, a.alpha)]
    with Pool(a.workers) as pool:
        res = pool.map(one_chain, jobs, chunksize=1)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f)
    plain = {r["chain"]: r for r in res if r["alpha"] == 0.0}
    inf = {r["chain"]: r for r in res if r["alpha"] == a.alpha}
    cs = sorted(plain)
    rng = np.random.default_rng(0)
    boot = [rng.integers(0, len(cs), len(cs)) for _ in range(4000)]
    print(f"{len(cs)} chains, nskip {a.nskip}, ndpost {a.ndpost}, alpha {a.alpha}")
    print("| per chain | plain (median) | informed (median) | ratio of sums over chains [95% bootstrap] |")
    print("|---|---|---|---|")
    for k, name in (("ess_sigma", "ESS(sigma)"), ("ess_f_p5", "ESS(f(x_i)), 5th percentile over rows"),
                    ("ess_f_median", "ESS(f(x_i)), median over rows"), ("ess_f_min", "ESS(f(x_i)), minimum over rows"),
                    ("min_ess", "min-ESS (min of the above)"), ("births", "accepted births"),
                    ("sec", "CPU seconds")):
        x = np.array([plain[c][k] for c in cs], float)
        y = np.array([inf[c][k] for c in cs], float)
        lo, hi = np.percentile([y[i].sum() / x[i].sum() for i in boot], [2.5, 97.5])
        print(f"| {name} | {np.median(x):.2f} | {np.median(y):.2f} | {y.sum() / x.sum():.3f} [{lo:.3f}, {hi:.3f}] |")
    cor = [np.corrcoef(plain[c]["post_mean"], inf[c]["post_mean"])[0, 1] for c in cs]
    print(f"posterior-mean correlation per chain: median {np.median(cor):.4f}, minimum {min(cor):.4f}")


if __name__ == "__main__":
    main()
