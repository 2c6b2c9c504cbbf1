"""Exact posterior over single-tree structures for tiny problems (test infrastructure).

For m = 1 tree the sampler's target over tree structures T is

    pi(T | y)  propto  prior(T) * integral p(y | T, sigma) p(sigma) dsigma

with the tree prior under which BART's birth/death ratios (src/bart_kernel.cpp:440-462) are exact:
  * an internal node at depth d with rule (v, c):  PG[d] * (1/p) * 1/(U_v - L_v + 1)   (L_v, U_v: the cutpoint range of
    v at the node, from its ancestors' rules -- src/bart_kernel.cpp:238-251);
  * a leaf at depth d:  1 - PG[d] if some variable has a non-empty range there (and d + 1 < 24), else 1;
restricted to trees whose every node holds at least 5 rows (births need both children to have 5 or more,
src/bart_kernel.cpp:485, and rows at a node do not change once its ancestors are fixed); and, given T and sigma, conjugate normal leaves mu ~ N(0, tau^2) with y_i ~ N(mu, sigma^2 / wt_i), sigma^2 ~
nu lambda / chi^2_nu. The integral over sigma is taken on a fine grid of log sigma.

Trees are represented canonically as nested tuples: None for a leaf, (v, c, left, right) for a split.
"""
from __future__ import annotations

import math

import numpy as np

MAXDEPTH = 24


class TreePrior:
    def __init__(self, numcut, pgtab):
        self.numcut = [int(x) for x in numcut]
        self.p = len(self.numcut)
        self.PG = [float(np.float32(x)) for x in pgtab]

    def trees(self):
        """[(tree, log prior)] over every tree the sampler can reach."""
        return list(self._enum(tuple((0, nc - 1) for nc in self.numcut), 0))

    def _enum(self, ranges, d):
        good = any(U >= L for (L, U) in ranges) and d + 1 < MAXDEPTH
        pg = self.PG[d] if good else 0.0
        if pg < 1.0:
            yield None, math.log1p(-pg)
        if not good:
            return
        for v, (L, U) in enumerate(ranges):
            if U < L:
                continue
            for c in range(L, U + 1):
                lr = list(ranges)
                lr[v] = (L, c - 1)
                rr = list(ranges)
                rr[v] = (c + 1, U)
                base = math.log(pg) - math.log(self.p) - math.log(U - L + 1)
                for lt, lp in self._enum(tuple(lr), d + 1):
                    for rt, rp in self._enum(tuple(rr), d + 1):
                        yield (v, c, lt, rt), base + lp + rp


def leaf_of(tree, q, i):
    node, path = tree, ()
    while node is not None:
        v, c, lt, rt = node
        if q[v][i] <= c:
            node, path = lt, path + (0,)
        else:
            node, path = rt, path + (1,)
    return path


def min_node_rows(tree, q, n):
    """Smallest row count over the children of every split (n for a root-only tree)."""
    best = n

    def walk(node, rows):
        nonlocal best
        if node is None:
            return
        v, c, lt, rt = node
        left = [i for i in rows if q[v][i] <= c]
        right = [i for i in rows if q[v][i] > c]
        best = min(best, len(left), len(right))
        walk(lt, left)
        walk(rt, right)

    walk(tree, list(range(n)))
    return best


def log_marginal(tree, q, y, wt, tau, nu, lam, log_sigma_grid):
    """log of integral p(y | T, sigma) p(sigma) over the grid (up to a constant shared by all trees)."""
    n = len(y)
    leaves = {}
    for i in range(n):
        leaves.setdefault(leaf_of(tree, q, i), []).append(i)
    t2 = float(tau) ** 2
    s = log_sigma_grid
    sig2 = np.exp(2 * s)
    ll = -n * s - 0.5 * np.sum(wt * y * y) / sig2
    for rows in leaves.values():
        b = np.sum(wt[rows]) / sig2
        M = np.sum(wt[rows] * y[rows]) / sig2
        k = 1.0 + t2 * b
        ll = ll - 0.5 * np.log(k) + 0.5 * t2 * M * M / k
    lp = -nu * s - 0.5 * nu * lam * np.exp(-2 * s)
    a = ll + lp
    mx = a.max()
    return float(mx + np.log(np.sum(np.exp(a - mx))))


def posterior(prob):
    """{tree: probability} for a problem dict (q [p][n], y, wt, numcut, pgtab, tau, nu, lambda)."""
    q = np.asarray(prob["q"]).astype(int)
    y = np.asarray(prob["y"], np.float64)
    wt = np.asarray(prob["wt"], np.float64)
    prior = TreePrior(prob["numcut"], prob["pgtab"])
    grid = np.linspace(math.log(1e-3), math.log(1e3), 6001)
    out = {}
    for tree, lp in prior.trees():
        if min_node_rows(tree, q, len(y)) < 5:
            continue
        out[tree] = lp + log_marginal(tree, q, y, wt, float(np.float32(prob["tau"])), float(prob["nu"]),
                                      float(prob["lambda"]), grid)
    mx = max(out.values())
    z = sum(math.exp(v - mx) for v in out.values())
    return {t: math.exp(v - mx) / z for t, v in out.items()}


def tree_from_words(words):
    """Canonical tree of one draw's 64 output words (v | c << 8 | l << 16 | r << 24 | theta << 32)."""
    w = np.asarray(words, np.uint64)

    def node(s):
        x = int(w[s])
        l, r = (x >> 16) & 255, (x >> 24) & 255
        if l == 0:
            return None
        return (x & 255, (x >> 8) & 255, node(l), node(r))

    return node(0)
