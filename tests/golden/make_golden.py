"""Write the golden trajectories of the reference sampler core.

Runs in the build container only (it needs the reference sources): builds oracle/_ref/libbart_ref.so from the
reference's src/bart_kernel.cpp (oracle/Makefile), runs bart_fit_cpu with the internal generator on small cases and
stores inputs and outputs as compressed .npz fixtures next to this script:

    python tests/golden/make_golden.py

Cases cover: n not a multiple of 16, a binary column, few cutpoints (so bad-variable proposals occur), weights far from
1, unit weights, p = 1, probit, a node cap hit (m = 1 deep trees), and the workload's own 16-predictor shape.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fastbart_amd import workloads as W  # noqa: E402
from oracle import pyoracle  # noqa: E402

CASES = [
    # name, problem builder, m, nskip, ndpost, seed, probit
    ("w37_p3", lambda: W.small_problem(37, 3, 5, 11, "uniform"), 5, 30, 10, 7, 0),
    ("w200_p3_extreme", lambda: W.small_problem(200, 3, 10, 12, "extreme"), 20, 40, 20, 3, 0),
    ("w1000_p16_ones", lambda: W.small_problem(1000, 16, 20, 13, "ones"), 20, 50, 20, 99, 0),
    ("w333_p1", lambda: W.small_problem(333, 1, 3, 15, "uniform"), 10, 40, 10, 123456789, 0),
    ("p300_probit", lambda: W.small_problem(300, 3, 10, 14, probit=True, ntree=10), 10, 30, 20, 5, 1),
    ("w517_m1_deep", lambda: W.small_problem(517, 4, 60, 16, "uniform", ntree=1), 1, 150, 20, 42, 0),
    ("w4096_p16_short", lambda: W.fwbart_problem(4096, 16, 100, 200, seed=0), 200, 3, 2, 1, 0),
]
# the same core built with 8 node slots: the node cap (first_free failing, stats_out[2]) is reached often
CASES_N8 = [
    ("n8_w800_cap", lambda: dict(W.small_problem(800, 4, 30, 21, "uniform", ntree=5), pgtab=W.pgtab(0.99, 0.5)),
     5, 80, 20, 9, 0),
]

SCALARS = ("tau", "lambda", "nu", "sigest", "offset")


def main() -> None:
    pyoracle.build(ref=True)
    refs = {64: pyoracle.RefCore(), 8: pyoracle.RefCore(pyoracle.REF_N8_SO, nodes=8)}
    for nodes, name, mk, m, nskip, nd, seed, probit in [(64,) + c for c in CASES] + [(8,) + c for c in CASES_N8]:
        ref = refs[nodes]
        pb = mk()
        out = ref.fit(pb, m, nskip, nd, seed, keep_trees=1, probit=probit)
        rec = dict(q=pb["q"], y=pb["y"], wt=pb["wt"], numcut=pb["numcut"], pgtab=pb["pgtab"],
                   ybin=pb["ybin"] if probit else np.zeros(0, np.int32),
                   scalars=np.array([pb[k] for k in SCALARS], np.float32),
                   run=np.array([m, nskip, nd, seed, probit, nodes], np.int64),
                   sigma=out["sigma"], trmean=out["trmean"], trees=out["trees"], stats=out["stats"])
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **rec)
        print(f"{name}: stats {out['stats'].tolist()} -> {os.path.getsize(path)} bytes")


def load(path: str) -> dict:
    """Fixture -> (problem dict, run parameters, expected outputs)."""
    z = np.load(path, allow_pickle=False)
    sc = z["scalars"]
    pb = dict(q=z["q"], y=z["y"], wt=z["wt"], numcut=z["numcut"], pgtab=z["pgtab"],
              **{k: float(v) for k, v in zip(SCALARS, sc)})
    m, nskip, nd, seed, probit, nodes = (int(v) for v in z["run"])
    if probit:
        pb["ybin"] = z["ybin"]
    exp = dict(sigma=z["sigma"], trmean=z["trmean"], trees=z["trees"], stats=z["stats"])
    return dict(problem=pb, m=m, nskip=nskip, ndpost=nd, seed=seed, probit=probit, nodes=nodes, expected=exp)


if __name__ == "__main__":
    main()
