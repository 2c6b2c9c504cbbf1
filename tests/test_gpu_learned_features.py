"""The proposal network's input features against their definition, on chains of the 512-chain workload.

For the node of a birth proposal, per variable v with cut range [L, U] at the node and each of 16 cut positions c
(L + 2^k - 1 and U - 2^k + 1, k = 0..7; positions outside [L, U] give 0), the feature is the split's standardized
log-likelihood gain z^2 / 2 - log(f (1 - f)) / 2, with f = W_c / W the left child's share of the node's weight and
z = (S_c - f S) / sqrt(W f (1 - f)) / sigma from the node's sums of w and w r (r the partial residual without the
tree); the last feature is the node's depth / 8. Here they are computed in NumPy from the batch state at the start
of a sweep (weights at the device's 2^-12 grid, w r at 2^-20) and compared with the pair the lock-step pipeline
records at the sweep's first tree update, for the chains whose first move (from the oracle's trace of the same chain)
is a birth."""
import math
import sys

import numpy as np
import pytest

from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.mlp import MLP
from test_gpu_split_target import _node_range, _tree_fields

pytestmark = pytest.mark.gpu

M = 200
PRE = 5  # plain sweeps before the learned sweep


def _feat_pos(L, U, k):
    if U < L:
        return -1
    c = L + (1 << k) - 1 if k < 8 else U - (1 << (15 - k)) + 1
    return c if L <= c <= U else -1


def expected_features(q, w, res, member, ranges, depth, sigma):
    p = q.shape[0]
    x = np.zeros(513)
    wf = np.rint(w[member] * 4096.0) / 4096.0
    wr = np.rint(w[member] * res[member] * 1048576.0) / 1048576.0
    Wt, St = wf.sum(), wr.sum()
    for v in range(p):
        L, U = ranges[v]
        qv = q[v][member]
        for k in range(16):
            c = _feat_pos(L, U, k)
            if c < 0:
                continue
            left = qv <= c
            f = min(max(wf[left].sum() / Wt, 1e-6), 1 - 1e-6)
            z = (wr[left].sum() - f * St) / math.sqrt(Wt * f * (1 - f)) / sigma
            x[16 * v + k] = 0.5 * z * z - 0.5 * math.log(f * (1 - f))
    x[512] = depth * 0.125
    return x


def test_gpu_learned_features_match_their_definition(oracle):
    net = MLP((513, 256, 256, 8160), max_batch=512, seed=5)
    checked = 0
    for chain in range(30):
        pb = W.batch512_chain(chain)
        # the oracle's trace says what the first tree update of sweep PRE proposes
        ch = oracle.chain(pb, M, 1000, 100, 1 + chain, trace_cap=M * (PRE + 1))
        ch.sweeps(PRE + 1)
        rec = [r for r in ch.trace() if r.it == PRE and r.j == 0][0]
        ch.close()
        if rec.mode != 1:
            continue
        with sampler.BatchSampler([pb], M, 1000, 100, [1 + chain]) as b:
            b.run(PRE)
            b.sync()
            st = b.state(0)
            b.learned_setup(net, 0.0, pair_capacity=1000)
            b.learned_target(1, seed=3)
            b.run_learned(1)
            got = b.learned_pairs_copy(0, 1)["features"][0]
        q = np.asarray(pb["q"], np.uint8)
        p = q.shape[0]
        v, c, l, r, th, par, depth, leaves = _tree_fields(st["trees"][0])
        nx = int(rec.nx)
        assert nx in leaves
        ranges = [_node_range(v, c, l, par, nx, var, int(pb["numcut"][var])) for var in range(p)]
        member = st["leaf"][0] == nx
        # a pair is recorded only where some rule leaves 5 rows on each side
        nrow = int(member.sum())
        if not any(5 <= int((q[u][member] <= cc).sum()) <= nrow - 5
                   for u in range(p) for cc in range(ranges[u][0], ranges[u][1] + 1)):
            continue
        res = st["latent"] - st["af"] + float(th[nx])
        want = expected_features(q, np.asarray(pb["wt"], np.float64), res, member, ranges, int(depth[nx]),
                                 float(st["sigma"]))
        err = np.abs(got - want)
        print(f"chain {chain}: node {nx}, {int(member.sum())} rows, max |feature| {np.abs(want).max():.3f}, "
              f"max error {err.max():.2e}", flush=True)
        sys.stdout.flush()
        assert np.all(err <= 1e-4 + 1e-5 * np.abs(want)), (chain, int(np.argmax(err)))
        assert np.all(got[16 * p:512] == 0)
        checked += 1
    assert checked >= 4
