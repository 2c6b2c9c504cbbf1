"""Brute-force pins of the oracle's sufficient statistics, independent of the reference core.

From the chain state at the start of a sweep (leaf ids, running fit, trees) the statistics of that sweep's first tree
update are recomputed here in numpy, by the rules of src/bart_kernel.cpp:164-212:
  * REF16: per-slot sums in 16 row lanes (row i in lane i % 16, rows in order, a lane starts at +0.0), the lanes then
    added in order 0..15 from +0.0;
  * FX64 (no reference counterpart): every addend rounded to the nearest multiple of 2^-32 (ties to even), summed
    exactly as int64, scaled back;
with r = (y - af) + theta[leaf] and the birth split left iff q[v] <= c. The log acceptance ratio is then rebuilt from
those sums with lh (src/bart_kernel.cpp:278) and compared to the oracle's. A dropped term, a wrong lane rule, a
transposed split test or a wrong fixed-point scale fails one of these."""
import math

import numpy as np
import pytest

from fastbart_amd import workloads as W

NODES = 64


def _decode(words):
    w = words.astype(np.uint64)
    v = (w & np.uint64(255)).astype(int)
    c = ((w >> np.uint64(8)) & np.uint64(255)).astype(int)
    l_ = ((w >> np.uint64(16)) & np.uint64(255)).astype(int)
    r_ = ((w >> np.uint64(24)) & np.uint64(255)).astype(int)
    th = (w >> np.uint64(32)).astype(np.uint32).view(np.float32).astype(np.float64)
    return v, c, l_, r_, th


def _ref16_sum(vals, mask):
    lanes = [0.0] * 16
    for i in np.nonzero(mask)[0]:
        lanes[i % 16] = lanes[i % 16] + float(vals[i])
    s = 0.0
    for k in range(16):
        s = s + lanes[k]
    return s


def _fx64_sum(vals, mask):
    img = np.rint(vals[mask] * 4294967296.0).astype(np.int64)
    return float(int(img.sum())) * (1.0 / 4294967296.0)


def _lh(b, M, t2):
    k = b * t2 + 1
    return -0.5 * math.log(k) + 0.5 * M * M * t2 / k


@pytest.mark.parametrize("stats_mode", [0, 1])
def test_first_update_statistics_by_brute_force(oracle, stats_mode):
    pb = W.small_problem(333, 3, 9, 71, "extreme", ntree=6)
    m, nskip, nd = 6, 30, 1
    ch = oracle.chain(pb, m, nskip, nd, 5, stats_mode=stats_mode, trace_cap=m * (nskip + nd))
    summ = _fx64_sum if stats_mode else _ref16_sum
    t2 = float(np.float32(pb["tau"])) ** 2
    checked_birth = checked_death = 0
    for k in range(nskip + nd - 1):
        st = ch.state()
        ntr = len(ch.trace())
        ch.sweeps(1)
        rec = ch.trace()[ntr]
        assert rec.it == k and rec.j == 0
        v, c, l_, r_, th = _decode(st["trees"][0])
        leaf = st["leaf"][0].astype(int)
        y = st["latent"]
        w = pb["wt"].astype(np.float64)
        r = (y - st["af"]) + th[leaf]
        wr = w * r
        for s in range(NODES):
            sel = leaf == s
            if not sel.any():
                continue
            assert summ(w, sel) == rec.S0[s], (k, s)
            assert summ(wr, sel) == rec.S1[s], (k, s)
        inv_s2 = 1.0 / (st["sigma"] * st["sigma"])
        if rec.mode == 1:
            left = (leaf == rec.nx) & (pb["q"][rec.v].astype(int) <= rec.c)
            right = (leaf == rec.nx) & ~(pb["q"][rec.v].astype(int) <= rec.c)
            assert (rec.nl, rec.nr) == (int(left.sum()), int(right.sum()))
            sl0, sl1, sr0, sr1 = summ(w, left), summ(wr, left), summ(w, right), summ(wr, right)
            assert (sl0, sl1, sr0, sr1) == (rec.sl0, rec.sl1, rec.sr0, rec.sr1), k
            if rec.nl >= 5 and rec.nr >= 5:
                bl, Ml, br, Mr = sl0 * inv_s2, sl1 * inv_s2, sr0 * inv_s2, sr1 * inv_s2
                la = math.log(rec.pr) + _lh(bl, Ml, t2) + _lh(br, Mr, t2) - _lh(bl + br, Ml + Mr, t2)
                assert math.isclose(min(la, 0.0), rec.lalpha, rel_tol=1e-12, abs_tol=1e-12), k
                checked_birth += 1
        elif rec.mode == 2:
            a, b = l_[rec.nx], r_[rec.nx]
            bl, Ml = summ(w, leaf == a) * inv_s2, summ(wr, leaf == a) * inv_s2
            br, Mr = summ(w, leaf == b) * inv_s2, summ(wr, leaf == b) * inv_s2
            la = math.log(rec.pr) + _lh(bl + br, Ml + Mr, t2) - _lh(bl, Ml, t2) - _lh(br, Mr, t2)
            assert math.isclose(min(la, 0.0), rec.lalpha, rel_tol=1e-12, abs_tol=1e-12), k
            checked_death += 1
    assert checked_birth > 0 and checked_death > 0


def test_fx64_statistics_are_order_free_and_total(oracle):
    """FX64: the leaf sums of w over all used leaves add up exactly to the fixed-point image of sum(w) (every row is in
    exactly one leaf), in any order of the rows."""
    pb = W.small_problem(517, 4, 12, 72, "uniform", ntree=5)
    ch = oracle.chain(pb, 5, 8, 1, 9, stats_mode=1, trace_cap=5 * 9)
    ch.sweeps(9)
    w = pb["wt"].astype(np.float64)
    total = int(np.rint(w * 4294967296.0).astype(np.int64).sum())
    for rec in ch.trace():
        img = [int(round(rec.S0[s] * 4294967296.0)) for s in range(NODES)]
        assert sum(img) == total
    perm = np.random.default_rng(0).permutation(len(w))
    assert _fx64_sum(w[perm], np.ones(len(w), bool)) == _fx64_sum(w, np.ones(len(w), bool))
