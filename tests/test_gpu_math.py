"""GPU arithmetic the sampler's bit-identity rests on: the float normal draw, correctly rounded sqrt and division, and
the agreement rate of the device double log/exp with the host libm (used only inside comparisons)."""
import ctypes as C

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler

pytestmark = pytest.mark.gpu


def test_gpu_sqrt_of_logf_image_exhaustive(oracle):
    """sqrtf(-2 logf(u)) on the device for all 2^24 grid points equals the host's."""
    k = (np.arange(1 << 24, dtype=np.uint64) << np.uint64(8)).astype(np.uint32)
    got = sampler.debug_math(1, u=k)
    assert same_bits(got, oracle.libm_image(3))


def test_gpu_normal_pairs(oracle):
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 2 ** 32, size=2 * 200000, dtype=np.uint64).astype(np.uint32)
    got = sampler.debug_math(0, u=raw)
    exp = oracle.normal_from(raw[0::2][:20000], raw[1::2][:20000])
    assert same_bits(got[:20000], exp)


def test_gpu_double_sqrt_and_division_are_correctly_rounded():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(0, 10, 200000), np.exp(rng.uniform(-700, 700, 200000)),
                        rng.uniform(1e-300, 1e-290, 1000)])
    assert same_bits(sampler.debug_math(4, d=x), np.sqrt(x))
    a = rng.standard_normal(400000) * np.exp(rng.uniform(-30, 30, 400000))
    b = rng.standard_normal(400000) * np.exp(rng.uniform(-30, 30, 400000))
    ab = np.empty(800000)
    ab[0::2], ab[1::2] = a, b
    assert same_bits(sampler.debug_math(5, d=ab), a / b)


def test_gpu_double_log_exp_agreement_with_host_libm():
    """log/exp on the device differ from glibc only in rare last-bit cases; the sampler uses them only in comparisons
    (lalpha against a log-uniform on a 2^24 grid, the gamma and truncated-normal acceptance tests), where a last-bit
    difference changes a decision only if the two sides are within an ulp."""
    libm = C.CDLL("libm.so.6")
    libm.log.argtypes = [C.c_double]
    libm.log.restype = C.c_double
    libm.exp.argtypes = [C.c_double]
    libm.exp.restype = C.c_double
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(1.0, 50.0, 50000), np.exp(rng.uniform(-20, 5, 50000))])
    gl = np.array([libm.log(float(v)) for v in x])
    dl = sampler.debug_math(2, d=x)
    e = -0.5 * rng.uniform(0, 8, 50000) ** 2
    ge = np.array([libm.exp(float(v)) for v in e])
    de = sampler.debug_math(3, d=e)
    mis_log = np.count_nonzero(gl != dl) / len(x)
    mis_exp = np.count_nonzero(ge != de) / len(e)
    print(f"log mismatch rate {mis_log:.2e}, exp mismatch rate {mis_exp:.2e}")
    # both are faithful (within one ulp of glibc); exp is not correctly rounded on the device (a few % last-bit
    # differences), which is harmless where it is used: u > exp(t) with u a float on the 2^-24 grid can only change
    # when u equals one of the two doubles exactly.
    assert np.max(np.abs(gl - dl) / np.abs(gl + 1e-300)) < 4e-16
    assert np.max(np.abs(ge - de) / ge) < 4e-16
    assert mis_log < 0.02 and mis_exp < 0.2


@pytest.mark.parametrize("case", ["fpbart_n24576", "batch512_chain0"])
def test_gpu_libm_decision_risk_on_full_workload_fits(oracle, case):
    """How often the device's double exp / log could change a decision of the reference's sampler, measured on every
    acceptance test of a whole fit as the oracle makes it (1,100 sweeps): rtnorm's exponential branch (u > exp(arg),
    src/bart_kernel.cpp:289) on the probit workload, the gamma draw (log u < A + d log v, :80) on a chain of the
    512-chain workload. Each logged argument is evaluated with the device's exp / log; a decision flips when the
    comparison's left side falls between the host and the device value. With the drawn uniforms no decision flips
    (and the comparisons that come within 1e-13 of a tie are all logged and checked); over random uniforms the
    expected number of flips per fit is the sum of the probabilities that a uniform lands between the two values,
    estimated on the first 200,000 comparisons and scaled to all of them."""
    from fastbart_amd import workloads as W
    if case == "fpbart_n24576":
        pb, m, probit = W.workload_fpbart_n24576(), 50, 1
    else:
        pb, m, probit = W.batch512_chain(0), 200, 0
    sample = 200_000
    recs, seen, _ = oracle.fit_logging_comparisons(pb, m, 1000, 100, 1, probit=probit, cap=2 * sample, sample=sample)
    kind, arg, lhs, A, d, host, near = (recs[:, k] for k in range(7))
    dev = np.empty(len(recs))
    e0, e1 = kind == 0, kind == 1
    if e0.any():
        dev[e0] = sampler.debug_math(3, d=np.ascontiguousarray(arg[e0]))
    if e1.any():
        dev[e1] = sampler.debug_math(2, d=np.ascontiguousarray(arg[e1]))
    # the decisions with the drawn uniforms: host value against device value
    flips0 = np.count_nonzero(e0 & ((lhs > host) != (lhs > dev)))
    rh, rd = A + d * host, A + d * dev
    flips1 = np.count_nonzero(e1 & ((lhs < rh) != (lhs < rd)))
    # expected flips for a uniform u on (0, 1): P(u between exp(arg)_host and exp(arg)_dev) = |difference|; for the
    # gamma test P(log u between the two right sides) = |exp(rh) - exp(rd)|
    ns = min(sample, len(recs))
    p0 = np.where(e0[:ns], np.abs(host[:ns] - dev[:ns]), 0.0)
    p1 = np.where(e1[:ns], np.abs(np.exp(rh[:ns]) - np.exp(rd[:ns])), 0.0)
    differ = np.count_nonzero(host[:ns] != dev[:ns]) / ns
    expected = (p0.sum() + p1.sum()) * seen / ns
    print(f"{case}: {seen} comparisons per fit, {int(near.sum())} within 1e-13 of a tie, device value differs from "
          f"the host's in {differ:.2%} of them; flips with the drawn uniforms {flips0 + flips1}; expected flips per "
          f"fit over random uniforms {expected:.3e}")
    assert flips0 + flips1 == 0
    assert expected < 1e-6
