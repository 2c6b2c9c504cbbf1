"""The learned birth proposal leaves the posterior unchanged: on a problem small enough to enumerate every tree, the
visit frequencies of single-tree chains match the exact posterior (tests/tree_posterior.py, itself checked against the
reference's sampler in test_tree_posterior.py) for the plain kernel, for the lock-step pipeline with BART's proposal
(alpha = 0), and for the pipeline with the network in the mixture (alpha = 0.8) -- including a network that puts
almost all of its mass on one rule, which without the Metropolis-Hastings correction would bias the chain towards it."""
import numpy as np
import pytest

import tree_posterior as TP
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd.mlp import MLP
from test_tree_posterior import check_against, tiny_problem

pytestmark = pytest.mark.gpu

CHAINS, NSKIP, NDPOST = 256, 60, 400


def _trees(b):
    out = []
    for c in range(CHAINS):
        tr = b.fetch(c)["trees"]
        out.append([TP.tree_from_words(tr[d, 0]) for d in range(NDPOST)])
    return out


def _net(kind):
    net = MLP((513, 64, 64, 8160), max_batch=CHAINS, seed=11)
    p = net.params()
    if kind == "peaked":
        p[:] = 0
        W1, b1, W2, b2, W3, b3 = net.unpack(p)
        b3[0 * 255 + 0] = 6.0        # rule (v = 0, c = 0): about 99% of the network's mass wherever it is allowed
    else:
        p *= 3.0                     # a random network with sharp, feature-dependent preferences
    net.set_params(p)
    return net


@pytest.fixture(scope="module")
def posterior():
    pb = tiny_problem()
    return pb, TP.posterior(pb)


def test_gpu_plain_kernel_matches_enumerated_posterior(posterior):
    pb, post = posterior
    with sampler.BatchSampler([pb] * CHAINS, 1, NSKIP, NDPOST, list(range(1, CHAINS + 1))) as b:
        b.run(NSKIP + NDPOST)
        tv, zmax, k = check_against(post, _trees(b))
    print(f"plain kernel: TV {tv:.4f} max|z| {zmax:.2f} over {k} trees")
    assert tv < 0.02 and zmax < 5.0


@pytest.mark.parametrize("alpha,kind", [(0.0, None), (0.8, "peaked"), (0.8, "random")])
def test_gpu_learned_pipeline_matches_enumerated_posterior(posterior, alpha, kind):
    pb, post = posterior
    net = _net(kind) if kind else None
    with sampler.BatchSampler([pb] * CHAINS, 1, NSKIP, NDPOST, list(range(1, CHAINS + 1))) as b:
        b.learned_setup(net, alpha)
        b.run_learned(NSKIP + NDPOST)
        tv, zmax, k = check_against(post, _trees(b))
    print(f"alpha {alpha} {kind}: TV {tv:.4f} max|z| {zmax:.2f} over {k} trees")
    assert tv < 0.02 and zmax < 5.0


def test_gpu_learned_pipeline_collects_pairs_and_trains():
    """Burn-in with BART's proposal records accepted births (features of the node, allowed rules, chosen rule); the
    network trained on them on the device fits them better than the uniform rule."""
    probs = [W.batch512_chain(c) for c in range(64)]
    net = MLP((513, 256, 256, 8160), max_batch=512, seed=5)
    with sampler.BatchSampler(probs, 200, 20, 0, list(range(1, 65)), keep_trees=0) as b:
        b.run(10)
        b.learned_setup(net, 0.0, pair_capacity=200000)
        b.run_learned(10)
        npairs = b.learned_pairs()
        losses = b.learned_train(300, batch=256, lr=1e-3, seed=2)
        st = [b.fetch(c)["stats"] for c in range(64)]
    print("pairs", npairs, "loss first", losses[:10].mean(), "last", losses[-10:].mean())
    assert npairs > 500
    assert np.isfinite(losses).all() and losses[-10:].mean() < losses[:10].mean()
    assert sum(s[0] for s in st) >= npairs
