"""The proposal network's conditional target on the GPU: the device's exact split conditional at a node against the
oracle's (orc_split_conditional, itself pinned by brute force in tests/test_split_conditional.py), on random nodes
and on the nodes of a real chain of the 512-chain workload; and the lock-step pipeline collecting pairs with that
target, which must leave the chains exactly as the oracle runs them."""
import sys

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.mlp import MLP
from test_split_conditional import CASES, pg_table, random_node

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(oracle, args, sigma, tau, u):
    q, member, w, r, ranges, depth, ngv, ngood, pg = args
    want, hw, nr = oracle.split_conditional(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau)
    got, h, cls = sampler.debug_split_conditional(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau, u)
    if nr == 0:
        assert h == -1.0 and not got.any()
        return 0
    # device: int64 fixed-point sums at 2^-20, log weights stored as float
    assert np.allclose(got, want, rtol=2e-4, atol=1e-9), float(np.abs(got - want).max())
    assert ((got > 0) == (want > 0)).all()
    assert abs(h - hw) < 1e-3 * max(1.0, abs(hw))
    # inverse CDF in class order v * 255 + c
    cum = np.cumsum(want.ravel())
    for t, k in enumerate(cls):
        assert want.ravel()[k] > 0
        lo = cum[k - 1] if k > 0 else 0.0
        assert lo - 1e-5 <= u[t] <= cum[k] + 1e-5, (t, k, u[t], lo, cum[k])
    return nr


@pytest.mark.parametrize("case", CASES)
def test_gpu_split_conditional_matches_oracle_random_nodes(oracle, case):
    n, p, ncut, depth, ngv, ngood, sigma, tau, signal = case
    _say("split conditional", case)
    rng = np.random.default_rng(n * 7 + p)
    q, member, w, r, ranges = random_node(rng, n, p, ncut, signal=signal)
    u = np.random.default_rng(5).random(40)
    nr = _compare(oracle, (q, member, w, r, ranges, depth, ngv, ngood, pg_table()), sigma, tau, u)
    assert nr > 0


def _tree_fields(words):
    w = np.asarray(words, np.uint64)
    v = (w & 0xFF).astype(int)
    c = ((w >> 8) & 0xFF).astype(int)
    l = ((w >> 16) & 0xFF).astype(int)
    r = ((w >> 24) & 0xFF).astype(int)
    th = (w >> 32).astype(np.uint32).view(np.float32)
    par = np.zeros(64, int)
    depth = np.zeros(64, int)
    leaves = []
    stack = [0]
    while stack:  # the slots reached from the root (unused slots keep stale fields)
        s = stack.pop()
        if l[s]:
            for ch in (l[s], r[s]):
                par[ch] = s
                depth[ch] = depth[s] + 1
                stack.append(ch)
        else:
            leaves.append(s)
    return v, c, l, r, th, par, depth, sorted(leaves)


def _node_range(v, c, l, par, node, var, nc):
    lo, hi = 0, nc - 1
    while node != 0:
        pa = par[node]
        if v[pa] == var:
            if l[pa] == node:
                hi = min(hi, c[pa] - 1)
            else:
                lo = max(lo, c[pa] + 1)
        node = pa
    return lo, hi


def test_gpu_split_conditional_matches_oracle_on_chain_nodes(oracle):
    """Every leaf of the first trees of two workload chains after 30 sweeps of the oracle: the node's rows, weights
    and partial residuals as the sampler sees them."""
    checked = 0
    for chain in (0, 5):
        pb = W.batch512_chain(chain)
        ch = oracle.chain(pb, 200, 30, 0, 1 + chain)
        ch.sweeps(30)
        st = ch.state()
        q = np.asarray(pb["q"], np.uint8)
        p, n = q.shape
        wts = np.asarray(pb["wt"], np.float64)
        sigma = float(st["sigma"])
        for j in range(0, 12):
            v, c, l, r, th, par, depth, leaves = _tree_fields(st["trees"][j])
            lf = st["leaf"][j]
            rng_nodes = {}
            for s in leaves:
                rg = [_node_range(v, c, l, par, s, var, int(pb["numcut"][var])) for var in range(p)]
                rng_nodes[s] = rg
            ngood = sum(1 for s in leaves if any(U >= L for L, U in rng_nodes[s]))
            for s in leaves:
                rg = rng_nodes[s]
                ranges = np.array([(L | (U << 8)) if U >= L else 0x00FF for L, U in rg], np.uint16)
                ngv = sum(1 for L, U in rg if U >= L)
                member = (lf == s).astype(np.uint8)
                res = st["latent"] - st["af"] + float(th[s])
                u = np.random.default_rng(j).random(8)
                checked += _compare(oracle, (q, member, wts, res, ranges, int(depth[s]), ngv, ngood, pg_table()),
                                    sigma, float(pb["tau"]), u) > 0
        ch.close()
    _say("chain nodes with rules:", checked)
    assert checked >= 20


def test_gpu_conditional_target_collects_pairs_and_leaves_the_chains_unchanged(oracle):
    """16 workload chains: 4 plain sweeps, then 4 lock-step sweeps at alpha 0 collecting every proposed birth's node
    with its split conditional as a soft target. The chains' outputs equal the oracle's plain fit bit for bit (the
    target's draws come from a stream apart from the chains' generators); the labels lie in their pairs' ranges, the
    conditionals' mean entropy lies below the uniform rule's level, and training on them lowers the loss on held-out
    pairs (the last fifth, from the latest sweeps)."""
    nch = 16
    probs = [W.batch512_chain(c) for c in range(nch)]
    seeds = [1 + c for c in range(nch)]
    m, nskip, nd = 200, 6, 2
    net = MLP((513, 256, 256, 8160), max_batch=512, seed=5)
    with sampler.BatchSampler(probs, m, nskip, nd, seeds) as b:
        b.run(4)
        b.learned_setup(net, 0.0, pair_capacity=100000)
        b.learned_target(1, seed=7)
        b.run_learned(nskip + nd - 4)
        outs = [b.fetch(c) for c in range(nch)]
        npairs = b.learned_pairs()
        ent, lall = b.learned_target_stats()
        pairs = b.learned_pairs_copy(0, 50)
        hold = npairs // 5
        b.learned_holdout(hold)
        before = b.learned_eval(npairs - hold, hold)
        losses = b.learned_train(300, batch=512, lr=1e-3, seed=3)
        after = b.learned_eval(npairs - hold, hold)
    _say("pairs", npairs, "mean entropy", ent, "mean log allowed", lall, "training loss", float(losses[:10].mean()),
         "->", float(losses[-10:].mean()), "held-out", before, "->", after)
    for c in range(nch):
        e = oracle.fit(probs[c], m, nskip, nd, seeds[c])
        for k in ("sigma", "trmean", "trees", "stats"):
            assert same_bits(outs[c][k], e[k]), (c, k)
    # about half the tree updates propose a birth; nodes with no qualifying rule give no pair
    assert 0.25 * nch * m * 4 < npairs <= 0.55 * nch * m * 4
    assert np.isfinite(ent) and np.isfinite(lall) and 0 < ent < lall
    # every label lies within its pair's allowed ranges
    for k in range(50):
        v, cut = divmod(int(pairs["cls"][k]), 255)
        r = int(pairs["ranges"][k, v])
        assert (r & 255) <= cut <= (r >> 8), k
    assert np.isfinite(losses).all()
    assert after[0] < before[0] and after[0] < after[2]
    assert after[1] < after[0]
