"""Oracle fits of workload chains in CPU worker processes, for GPU tests whose oracle side takes seconds per chain
(test infrastructure; imports nothing that touches the GPU)."""
from fastbart_amd import workloads as W

KEYS = ("sigma", "trmean", "trees", "stats")


def fit_batch512_chain(args):
    """chain c of the 512-chain workload, m trees, nskip + ndpost sweeps, seed 1 + c, through the oracle"""
    c, m, nskip, ndpost = args
    from oracle import pyoracle
    e = pyoracle.Oracle().fit(W.batch512_chain(c), m, nskip, ndpost, 1 + c)
    return c, {k: e[k] for k in KEYS}
