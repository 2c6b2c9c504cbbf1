"""The C restatement against the reference core's own outputs (fixtures written by tests/golden/make_golden.py from
the compiled reference src/bart_kernel.cpp, internal generator): every output bit-identical."""
import numpy as np
import pytest

from conftest import same_bits


CASES = ["w37_p3", "w200_p3_extreme", "w1000_p16_ones", "w333_p1", "p300_probit", "w517_m1_deep",
         "w4096_p16_short", "n8_w800_cap"]


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_trajectory(oracle, golden_cases, name):
    g = golden_cases[name]
    out = oracle.fit(g["problem"], g["m"], g["nskip"], g["ndpost"], g["seed"], probit=g["probit"],
                     node_cap=g["nodes"])
    e = g["expected"]
    assert same_bits(out["sigma"], e["sigma"])
    assert same_bits(out["trmean"], e["trmean"])
    assert np.array_equal(out["stats"], e["stats"])
    nodes = g["nodes"]
    assert same_bits(out["trees"][:, :, :nodes], e["trees"])
    # slots beyond the reference's NODES are never used: unused-slot words (v = 255, zero fields)
    assert np.all(out["trees"][:, :, nodes:] == 255)


def test_golden_cases_exercise_every_path(golden_cases):
    """The fixtures between them hit births, deaths, probit, p = 1, unit and extreme weights and the node cap."""
    assert any(g["probit"] for g in golden_cases.values())
    assert any(g["problem"]["q"].shape[0] == 1 for g in golden_cases.values())
    assert any(g["expected"]["stats"][2] > 0 for g in golden_cases.values())
    assert all(g["expected"]["stats"][0] > 0 and g["expected"]["stats"][1] > 0 for g in golden_cases.values())
    assert any(g["problem"]["q"].shape[1] % 16 for g in golden_cases.values())


def test_oracle_chain_resume_is_bit_identical(oracle, golden_cases):
    """k sweeps, then the rest, equals one uninterrupted run (checkpoint/resume of the chain state)."""
    g = golden_cases["w200_p3_extreme"]
    pb, m, nskip, nd, seed = g["problem"], g["m"], g["nskip"], g["ndpost"], g["seed"]
    ch = oracle.chain(pb, m, nskip, nd, seed)
    ch.sweeps(17)
    ch.sweeps(nskip + nd - 17)
    assert same_bits(ch.sigma_out, g["expected"]["sigma"])
    assert same_bits(ch.trmean(), g["expected"]["trmean"])
    assert same_bits(ch.trees_out.reshape(nd, m, 64), g["expected"]["trees"])


def test_oracle_is_reentrant(oracle, golden_cases):
    """Two chains interleaved sweep by sweep give the same results as each alone (the reference core cannot)."""
    a, b = golden_cases["w37_p3"], golden_cases["w333_p1"]
    ca = oracle.chain(a["problem"], a["m"], a["nskip"], a["ndpost"], a["seed"])
    cb = oracle.chain(b["problem"], b["m"], b["nskip"], b["ndpost"], b["seed"])
    for _ in range(60):
        ca.sweeps(1)
        cb.sweeps(1)
    assert same_bits(ca.sigma_out, a["expected"]["sigma"])
    assert same_bits(cb.sigma_out, b["expected"]["sigma"])
