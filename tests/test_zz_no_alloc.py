"""The allocation-free prediction entry point (fbh_predict_rows_device_noalloc) against the entry point it varies
(fbh_predict_rows_device), bit for bit, through the C-ABI: two calls in a row, the second with a smaller input, in both
kernel forms; and its size limit."""
import sys

import numpy as np
import pytest

from conftest import same_bits
from fastbart_amd import sampler
from fastbart_amd import workloads as W
from fastbart_amd._lib import FbhError, check, lib
from fastbart_amd.predict import CompiledTrees, DeviceBuffer, DeviceRows, device_sync

pytestmark = pytest.mark.gpu

SCRATCH_RANKS = 8 << 20


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc():
    return lib().fbh_predict_rows_device_noalloc


@pytest.fixture(scope="module")
def fit():
    pb = W.batch512_chain(3)
    with sampler.BatchSampler([pb], 200, 30, 40, [4], keep_trees=1) as b:
        b.run(70)
        out = b.fetch(0)
    text, _ = sampler.tree_text(out["trees"], pb["q"].shape[0])
    return pb, text


def _both(ct, x, draws):
    """(existing entry point, allocation-free entry point) outputs for rows x and the 1-based draw list."""
    rows = DeviceRows(x)
    d_draws = DeviceBuffer.from_array(np.asarray(draws, np.int32))
    ref = DeviceBuffer(8 * len(draws) * rows.n)
    got = DeviceBuffer(8 * len(draws) * rows.n)
    # the output the allocation-free call must overwrite starts as NaN
    got_init = np.full((len(draws), rows.n), np.nan)
    check(lib().fbh_device_copy(got.ptr, got_init.ctypes.data, got_init.nbytes, 0, 0), "fbh_device_copy")
    rows.predict_device(ct, d_draws, len(draws), ref)
    device_sync()
    check(_noalloc()(ct.h, rows.h, d_draws.ptr, len(draws), ct.mu, got.ptr, None), "fbh_predict_rows_device_noalloc")
    device_sync()
    form = ct.last_launch()
    a = ref.to_array(np.float64, (len(draws), rows.n))
    b = got.to_array(np.float64, (len(draws), rows.n))
    for o in (ref, got, d_draws, rows):
        o.close()
    return a, b, form


def test_noalloc_matches_the_device_entry_point_rank_form(fit):
    pb, text = fit
    p = pb["q"].shape[0]
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.75)
    assert ct.rank_form
    rng = np.random.Generator(np.random.PCG64(11))
    # first call: 40 draws x 9,001 rows (two rows per thread); second: 3 draws x 1,337 rows (one row per thread)
    for n, draws, rpt in ((9001, np.arange(1, 41), 2), (1337, [40, 2, 17], 1)):
        x = rng.uniform(-0.05, 1.05, size=(n, p))
        _say("rank form", n, "rows", len(draws), "draws")
        a, b, form = _both(ct, x, draws)
        assert form == (1, rpt)
        assert np.isfinite(a).all()
        assert same_bits(a, b)
    ct.close()


def test_noalloc_matches_the_device_entry_point_general_form(fit):
    pb, text = fit
    p = pb["q"].shape[0]
    rev = [np.asarray(c)[::-1].copy() for c in pb["cutpoints"]]
    ct = CompiledTrees(text, rev, mu=-0.5)
    assert not ct.rank_form
    rng = np.random.Generator(np.random.PCG64(12))
    for n, draws in ((3000, np.arange(1, 21)), (700, [5, 9])):
        x = rng.uniform(0, 1, size=(n, p))
        a, b, form = _both(ct, x, draws)
        assert form == (0, 1)
        assert same_bits(a, b)
    ct.close()


def test_noalloc_rejects_rows_beyond_the_scratch(fit):
    pb, text = fit
    p = pb["q"].shape[0]
    ct = CompiledTrees(text, pb["cutpoints"], mu=0.0)
    n = SCRATCH_RANKS // p + 1
    rows = DeviceRows(np.zeros((n, p)))
    d_draws = DeviceBuffer.from_array(np.array([1], np.int32))
    out = DeviceBuffer(8 * n)
    rc = _noalloc()(ct.h, rows.h, d_draws.ptr, 1, 0.0, out.ptr, None)
    assert rc == -1
    assert b"exceeds the scratch" in lib().fbh_last_error()
    # at the limit it runs
    rows2 = DeviceRows(np.zeros((SCRATCH_RANKS // p, p)))
    out2 = DeviceBuffer(8 * rows2.n)
    check(_noalloc()(ct.h, rows2.h, d_draws.ptr, 1, 0.0, out2.ptr, None), "noalloc at the limit")
    device_sync()
    for o in (rows, rows2, d_draws, out, out2):
        o.close()
    ct.close()
