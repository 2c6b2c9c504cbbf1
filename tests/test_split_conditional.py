"""The exact split conditional at a birth node (the training target of the proposal network; no reference
counterpart) in the oracle, against a brute-force restatement: every rule's children formed by masking the rows
directly, not by the prefix sums over per-rank histograms the oracle and the device use.

q(v, c) is proportional to q_unif(v, c) times the factors of BART's birth ratio that depend on the rule
(src/bart_kernel.cpp:441-451: the children's split probabilities 1 - PG, and PD, which depends on them when the tree
has one good leaf; :486: the children's likelihoods lh, :278), over the rules whose children hold at least 5 rows."""
import math

import numpy as np
import pytest


def _lh(b, M, t2):
    k = b * t2 + 1
    return -0.5 * math.log(k) + 0.5 * M * M * t2 / k


def brute(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau):
    p, n = q.shape
    inv_s2, t2 = 1.0 / sigma ** 2, tau ** 2
    idx = np.nonzero(member)[0]
    lw = {}
    for v in range(p):
        L, U = ranges[v] & 255, ranges[v] >> 8
        if U < L:
            continue
        for c in range(L, U + 1):
            left = idx[q[v, idx] <= c]
            right = idx[q[v, idx] > c]
            if len(left) < 5 or len(right) < 5:
                continue
            # the children may split again (PG at depth + 1) unless this was the node's only good variable and the
            # child's range on it is empty
            if ngv > 1:
                pgl = pgr = pg[depth + 1]
            else:
                pgl = 0.0 if c - 1 < L else pg[depth + 1]
                pgr = 0.0 if U < c + 1 else pg[depth + 1]
            pd = 0.5 if ngood > 1 else (1.0 if (pgl == 0 and pgr == 0) else 0.5)
            s0l, s1l = math.fsum(w[left]), math.fsum(w[left] * r[left])
            s0r, s1r = math.fsum(w[right]), math.fsum(w[right] * r[right])
            lw[(v, c)] = (math.log((1 - pgl) * (1 - pgr) * pd / (U - L + 1)) + _lh(s0l * inv_s2, s1l * inv_s2, t2)
                          + _lh(s0r * inv_s2, s1r * inv_s2, t2))
    out = np.zeros((p, 255))
    if not lw:
        return out, -1.0
    mx = max(lw.values())
    z = sum(math.exp(x - mx) for x in lw.values())
    for (v, c), x in lw.items():
        out[v, c] = math.exp(x - mx) / z
    qq = out[out > 0]
    return out, float(-(qq * np.log(qq)).sum())


def random_node(rng, n, p, ncut, frac=0.6, signal=2.0):
    q = rng.integers(0, ncut, size=(p, n)).astype(np.uint8)
    member = (rng.random(n) < frac).astype(np.uint8)
    w = rng.uniform(0.5, 2.0, n)
    r = rng.normal(size=n) + signal * (q[1 % p] > ncut // 2)  # a step on variable 1: a peaked conditional
    L = rng.integers(0, ncut // 3, size=p)
    U = rng.integers(ncut // 2, ncut, size=p)
    ranges = (L | (U << 8)).astype(np.uint16)
    return q, member, w, r, ranges


def pg_table(base=0.95, power=2.0):
    return np.array([base / (1 + d) ** power for d in range(64)])


CASES = [  # (n, p, ncut, depth, ngv, ngood, sigma, tau, signal)
    (200, 3, 40, 0, 3, 1, 1.0, 0.3, 2.0),
    (400, 5, 100, 2, 5, 4, 0.7, 0.1, 1.0),
    (300, 4, 255, 1, 1, 1, 1.5, 0.5, 0.0),   # one good variable: the children's PG at the range edges
    (150, 2, 12, 3, 1, 3, 1.0, 0.2, 3.0),
    (500, 8, 60, 5, 8, 2, 2.0, 0.05, 0.5),
]


@pytest.mark.parametrize("case", CASES)
def test_oracle_split_conditional_matches_brute_force(oracle, case):
    n, p, ncut, depth, ngv, ngood, sigma, tau, signal = case
    rng = np.random.default_rng(n * 7 + p)
    q, member, w, r, ranges = random_node(rng, n, p, ncut, signal=signal)
    if p > 2:
        ranges[2] = 0x00FF  # a variable with no cut left at the node
    pg = pg_table()
    got, h, nr = oracle.split_conditional(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau)
    want, hw = brute(q, member, w, r, ranges, depth, ngv, ngood, pg, sigma, tau)
    assert nr == int((want > 0).sum()) > 0
    assert np.allclose(got, want, rtol=1e-9, atol=1e-15)
    assert abs(h - hw) < 1e-9
    assert abs(got.sum() - 1) < 1e-12


def test_oracle_split_conditional_edge_cases(oracle):
    rng = np.random.default_rng(3)
    q, member, w, r, ranges = random_node(rng, 60, 3, 20)
    pg = pg_table()
    # fewer than 10 rows in the node: no rule has 5 rows on each side
    m9 = np.zeros(60, np.uint8)
    m9[:9] = 1
    got, h, nr = oracle.split_conditional(q, m9, w, r, ranges, 0, 3, 1, pg, 1.0, 0.3)
    assert nr == 0 and h == -1 and not got.any()
    # every range empty
    got, h, nr = oracle.split_conditional(q, member, w, r, np.full(3, 0x00FF, np.uint16), 0, 3, 1, pg, 1.0, 0.3)
    assert nr == 0 and h == -1
    # a flat residual and equal weights with one rank per row group: the likelihood terms still differ by child size,
    # but a single allowed cut gives probability one
    one = np.array([5 | (5 << 8), 0x00FF, 0x00FF], np.uint16)
    q1 = np.tile(np.repeat(np.arange(20, dtype=np.uint8), 3), (3, 1))
    got, h, nr = oracle.split_conditional(q1, np.ones(60, np.uint8), np.ones(60), np.zeros(60), one, 0, 3, 2, pg,
                                          1.0, 0.3)
    assert nr == 1 and got[0, 5] == 1.0 and abs(h) < 1e-15


def test_oracle_split_conditional_is_the_rule_posterior_of_a_birth(oracle):
    """With a flat prior over the depth terms (PG the same at every depth is not needed: with ngv > 1 and ngood > 1
    the prior factors are constant over rules), q is proportional to q_unif times the children's marginal
    likelihoods: the ratio of two rules' probabilities equals the ratio of their birth posteriors computed from the
    Gaussian marginal likelihood of each child written out in full (prior mean 0, variance tau^2, noise sigma^2/w)."""
    rng = np.random.default_rng(11)
    n, p, ncut = 80, 2, 10
    q, member, w, r, ranges = random_node(rng, n, p, ncut, frac=1.0, signal=1.5)
    ranges[:] = (0 | (9 << 8))
    sigma, tau = 0.8, 0.4
    got, _, _ = oracle.split_conditional(q, member, w, r, ranges, 1, 2, 2, pg_table(), sigma, tau)

    def log_marg(rows):
        # y_i = mu + e_i, e_i ~ N(0, sigma^2 / w_i), mu ~ N(0, tau^2): log p(y) up to terms equal for every split
        # of the same rows (the product of the per-row normal constants and the exponent's sum of w y^2)
        cov = np.diag(sigma ** 2 / w[rows]) + tau ** 2
        sign, logdet = np.linalg.slogdet(cov)
        y = r[rows]
        return -0.5 * logdet - 0.5 * y @ np.linalg.solve(cov, y) + 0.5 * np.sum(np.log(sigma ** 2 / w[rows])) + \
            0.5 * np.sum(w[rows] * y * y) / sigma ** 2

    lp = {}
    for v in range(p):
        for c in range(10):
            left, right = np.nonzero(q[v] <= c)[0], np.nonzero(q[v] > c)[0]
            if len(left) >= 5 and len(right) >= 5:
                lp[(v, c)] = log_marg(left) + log_marg(right)
    keys = list(lp)
    k0 = keys[0]
    for k in keys[1:]:
        assert got[k] > 0 and got[k0] > 0
        assert abs(math.log(got[k] / got[k0]) - (lp[k] - lp[k0])) < 1e-8, k
