"""The exact single-tree posterior (tests/tree_posterior.py) against a closed form and against long runs of the oracle
(the reference's sampler, m = 1): the enumeration is the target the learned proposal's exactness test uses."""
import math

import numpy as np

import tree_posterior as TP
from fastbart_amd import workloads as W


def tiny_problem(seed=3, n=40):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(0, 1, n)
    x1 = (rng.uniform(0, 1, n) < 0.5).astype(float)
    X = np.column_stack([x0, x1])
    # a weak signal, so that every tree with enough rows per node keeps some posterior mass and the chains mix
    y = 0.25 * (x0 > 0.55) - 0.2 * x1 + 0.5 * rng.standard_normal(n)
    xi = [np.array([0.3, 0.6]), np.array([0.5])]
    return W.wbart_problem(X, y, None, ntree=1, numcut=2, cutpoints=xi)


def chain_frequencies(trees_per_chain, support):
    """[chains][trees of the support] visit frequencies"""
    idx = {t: k for k, t in enumerate(support)}
    F = np.zeros((len(trees_per_chain), len(support)))
    for c, ts in enumerate(trees_per_chain):
        for t in ts:
            F[c, idx[t]] += 1
        F[c] /= max(len(ts), 1)
    return F


def check_against(post, trees_per_chain, zmax=5.0):
    support = sorted(post, key=lambda t: -post[t])
    F = chain_frequencies(trees_per_chain, support)
    mean, se = F.mean(axis=0), F.std(axis=0, ddof=1) / math.sqrt(F.shape[0])
    pi = np.array([post[t] for t in support])
    tv = 0.5 * np.abs(mean - pi).sum()
    big = pi > 0.005
    z = (mean[big] - pi[big]) / np.maximum(se[big], 1e-4)
    return tv, float(np.abs(z).max()), int(big.sum())


def test_prior_closed_form_one_variable_one_cut():
    pr = TP.TreePrior([1], np.array([0.95 / (1 + d) ** 2 for d in range(64)], np.float32))
    trees = dict(pr.trees())
    pg0 = float(np.float32(0.95))
    assert set(trees) == {None, (0, 0, None, None)}
    assert math.isclose(math.exp(trees[None]), 1 - pg0, rel_tol=1e-12)
    assert math.isclose(math.exp(trees[(0, 0, None, None)]), pg0, rel_tol=1e-12)   # children cannot split


def test_oracle_plain_sampler_matches_enumerated_posterior(oracle):
    pb = tiny_problem()
    post = TP.posterior(pb)
    assert abs(sum(post.values()) - 1) < 1e-9 and len(post) > 10
    per_chain = []
    for c in range(48):
        out = oracle.fit(pb, 1, 20, 600, 100 + c)
        per_chain.append([TP.tree_from_words(out["trees"][d, 0]) for d in range(600)])
    tv, zmax, k = check_against(post, per_chain)
    print(f"trees {len(post)}, checked {k}, TV {tv:.4f}, max |z| {zmax:.2f}")
    assert tv < 0.03 and zmax < 5.0


def test_oracle_informed_rule_keeps_the_enumerated_posterior(oracle):
    """The proposal study's informed birth rule (the rule's exact conditional mixed with the reference's uniform rule,
    its probability in the births' and deaths' ratios) leaves the target unchanged: the visit frequencies over tree
    structures match the enumeration, at alpha 0.8 and with the informed rule alone (alpha 1)."""
    pb = tiny_problem()
    post = TP.posterior(pb)
    for alpha in (0.8, 1.0):
        per_chain, informed = [], 0
        for c in range(48):
            ch = oracle.chain(pb, 1, 20, 600, 300 + c)
            ch.set_informed(alpha)
            ch.sweeps(620)
            informed += ch.informed_births
            t = ch.trees_out.reshape(600, 1, -1)
            per_chain.append([TP.tree_from_words(t[d, 0]) for d in range(600)])
            ch.close()
        tv, zmax, k = check_against(post, per_chain)
        print(f"alpha {alpha}: informed births {informed}, trees {len(post)}, checked {k}, TV {tv:.4f}, "
              f"max |z| {zmax:.2f}")
        assert informed > 1000
        assert tv < 0.03 and zmax < 5.0
