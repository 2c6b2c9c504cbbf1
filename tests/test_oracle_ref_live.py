"""The C restatement against the compiled reference core directly, on inputs beyond the committed fixtures
(runs where oracle/_ref was built; skipped otherwise)."""
import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import workloads as W
from oracle import pyoracle

pytestmark = pytest.mark.skipif(not pyoracle.RefCore.available(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def ref():
    return pyoracle.RefCore()


@pytest.mark.parametrize("seed", [1, 2, 77])
def test_batch_chain_shapes_match_reference(oracle, ref, seed):
    pb = W.batch512_chain(seed)
    a = oracle.fit(pb, 200, 4, 2, seed + 1)
    b = ref.fit(pb, 200, 4, 2, seed + 1)
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(a[k], b[k]), k


@pytest.mark.parametrize("n,seed", [(16, 3), (17, 4), (1500, 5)])
def test_probit_matches_reference(oracle, ref, n, seed):
    pb = W.small_problem(n, 3, 8, seed, probit=True, ntree=8)
    a = oracle.fit(pb, 8, 25, 10, seed, probit=1)
    b = ref.fit(pb, 8, 25, 10, seed, probit=1)
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(a[k], b[k]), k


def test_constant_columns_match_reference(oracle, ref):
    """No variable has a cutpoint: the root can never split (PBx = 0 -> death proposals on a root-only tree)."""
    pb = W.small_problem(64, 2, 4, 9)
    pb["numcut"] = np.zeros(2, np.int32)
    pb["q"] = np.zeros_like(pb["q"])
    a = oracle.fit(pb, 3, 10, 5, 5)
    b = ref.fit(pb, 3, 10, 5, 5)
    for k in ("sigma", "trmean", "trees", "stats"):
        assert same_bits(a[k], b[k]), k
