"""pansim_device_philox_noalloc (the entry point that allocates nothing)
against pansim_device_philox, through the C ABI: bit-equal blocks on two
calls in a row, the second with fewer counters (the shared scratch keeps the
first call's contents), and a clear error past the scratch size."""
import ctypes

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import philox

pytestmark = pytest.mark.gpu

MAX = 65536


def _lib():
    L = pansim_amd.lib()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.pansim_device_philox_noalloc.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t]
    return L


def _call(fn, ctr, k0, k1):
    out = np.zeros_like(ctr)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    rc = fn(ctr.ctypes.data_as(u32p), k0, k1, out.ctypes.data_as(u32p), ctr.shape[0])
    return rc, out


def test_noalloc_philox_matches_the_allocating_entry_point():
    L = _lib()
    rng = np.random.default_rng(7)
    for n, key in ((50_000, (0x1234, 0xBEEF)), (777, (5, 9))):
        ctr = rng.integers(0, 2 ** 32, size=(n, 4), dtype=np.uint32)
        rc_a, ref = _call(L.pansim_device_philox, ctr, *key)
        rc_b, got = _call(L.pansim_device_philox_noalloc, ctr, *key)
        assert rc_a == 0 and rc_b == 0
        assert np.array_equal(got, ref), n
        for i in (0, n // 2, n - 1):  # and the known-answer-pinned CPU Philox
            assert got[i].tolist() == philox(ctr[i].tolist(), key)
    too_many = np.zeros((MAX + 1, 4), dtype=np.uint32)
    rc, _ = _call(L.pansim_device_philox_noalloc, too_many, 0, 0)
    assert rc == -1
    L.pansim_last_error.restype = ctypes.c_char_p
    assert b"scratch" in L.pansim_last_error(None)
