"""The device's exp, log and Philox agree bit for bit with the oracle's (and so
with the known answers the oracle is pinned to): every random draw and every
infection ratio of the GPU path depends on them."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import pansim_amd
from oracle import oracle as O
from tests.test_oracle_math import PHILOX_KAT

pytestmark = pytest.mark.gpu


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def test_device_expf_matches_oracle_bitwise():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(-104.0, 89.0, 30000), -np.exp(rng.uniform(-28.0, 4.6, 30000)),
                         [0.0, -0.0, 88.7, -103.9, np.inf, -np.inf, np.nan]]).astype(np.float32)
    out = np.zeros_like(xs)
    assert pansim_amd.lib().pansim_device_expf(_f32p(xs), _f32p(out), xs.size) == 0
    ref = np.array([O.expf(float(x)) for x in xs], dtype=np.float32)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_device_log_matches_oracle_bitwise():
    rng = np.random.default_rng(2)
    xs = np.concatenate([np.exp(rng.uniform(-700.0, 700.0, 20000)), 1.0 - rng.uniform(0.0, 1.0, 20000),
                         [1.0, 0.5, 2.0, 5e-324, 0.0]]).astype(np.float64)
    out = np.zeros_like(xs)
    assert pansim_amd.lib().pansim_device_log(_f64p(xs), _f64p(out), xs.size) == 0
    ref = np.array([O.log(float(x)) for x in xs])
    assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_device_philox_known_answers(ctr, key, expected):
    c = np.array(ctr, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    assert pansim_amd.lib().pansim_device_philox(_u32p(c), key[0], key[1], _u32p(out), 1) == 0
    assert tuple(out.tolist()) == expected


def test_device_philox_matches_oracle():
    rng = np.random.default_rng(3)
    n = 4096
    c = rng.integers(0, 2 ** 32, size=(n, 4), dtype=np.uint64).astype(np.uint32)
    out = np.zeros((n, 4), dtype=np.uint32)
    assert pansim_amd.lib().pansim_device_philox(_u32p(c), 0x12345678, 0x9abcdef0, _u32p(out), n) == 0
    for i in range(0, n, 97):
        assert out[i].tolist() == O.philox(tuple(int(x) for x in c[i]), (0x12345678, 0x9abcdef0))
