"""Host side of the device-wide probit redraw: the jump tables (powers T^(128 2^k) of the generator's state
transition over GF(2)^128) that let every generating thread start at its own slot, against plain stepping of
xoshiro128** (src/bart_kernel.cpp:54-58)."""
import ctypes as C

import numpy as np
import pytest

from fastbart_amd._lib import lib

M32 = 0xFFFFFFFF


def _step(s):
    s0, s1, s2, s3 = s
    t = (s1 << 9) & M32
    s2 ^= s0
    s3 ^= s1
    s1 ^= s2
    s0 ^= s3
    s2 ^= t
    s3 = ((s3 << 11) | (s3 >> 21)) & M32
    return [s0, s1, s2, s3]


def _jump(state, blocks):
    st = np.array(state, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    rc = lib().fbh_debug_jump(st.ctypes.data_as(C.c_void_p), C.c_ulonglong(blocks), out.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return [int(x) for x in out]


@pytest.mark.parametrize("blocks", [0, 1, 2, 3, 5, 8, 37, 255, 1000])
def test_jump_tables_match_stepping(blocks):
    rng = np.random.default_rng(blocks + 1)
    state = [int(x) for x in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
    s = list(state)
    for _ in range(128 * blocks):
        s = _step(s)
    assert _jump(state, blocks) == s


def test_jump_is_linear_and_composes():
    """T^a T^b = T^(a + b), and the map is linear over GF(2): jump(x ^ y) = jump(x) ^ jump(y) -- a dropped or
    transposed table entry breaks one of them at large block counts that stepping would take too long to reach."""
    rng = np.random.default_rng(7)
    for _ in range(4):
        x = [int(v) for v in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
        y = [int(v) for v in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
        a, b = int(rng.integers(1, 1 << 19)), int(rng.integers(1, 1 << 19))
        assert _jump(_jump(x, a), b) == _jump(x, a + b)
        jx, jy = _jump(x, a), _jump(y, a)
        assert _jump([u ^ v for u, v in zip(x, y)], a) == [u ^ v for u, v in zip(jx, jy)]


def test_jump_rejects_out_of_range():
    st = np.zeros(4, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    assert lib().fbh_debug_jump(st.ctypes.data_as(C.c_void_p), C.c_ulonglong(1 << 20),
                                out.ctypes.data_as(C.c_void_p)) != 0
