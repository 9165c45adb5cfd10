"""The C-ABI library loads on a machine without a GPU, exports every entry
point that include/*.h declares, and fails loudly (no CPU fallback) when a
simulation is created without a HIP device; input errors are reported
before any device call (CPU only)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

import pansim_amd
from tests.conftest import ROOT, gpu_available

HEADERS = [os.path.join(ROOT, "include", h) for h in ("pansim_hip.h", "pansim_sim.h")]
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\b(pansim_[a-z0-9_]+)\s*\(", re.M)


def declared_symbols():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += DECL.findall(text)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(pansim_amd.LIB_PATH):
        from pansim_amd import build
        build.build_engine()
    return ctypes.CDLL(pansim_amd.LIB_PATH)


def test_headers_declare_the_engine_and_loop():
    names = declared_symbols()
    for required in ("pansim_create", "pansim_destroy", "pansim_movement_step", "pansim_infection_step",
                     "pansim_plan_locations", "pansim_progression_day", "pansim_testing_day", "pansim_stats_day",
                     "pansim_rebuild_locations", "pansim_sim_create", "pansim_sim_run_for_day"):
        assert required in names
    assert len(names) > 40


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"declared in include/ but not exported: {missing}"


def _create(lib, args):
    lib.pansim_sim_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                      ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    lib.pansim_sim_create_error.restype = ctypes.c_char_p
    args = pansim_amd.with_data_dir(args)
    argv = (ctypes.c_char_p * len(args))(*[a.encode() for a in args])
    h = ctypes.c_void_p()
    rc = lib.pansim_sim_create(ctypes.byref(h), len(args), argv, 0)
    return rc, lib.pansim_sim_create_error().decode(), h


def test_input_errors_are_reported_before_any_device_call(lib):
    rc, msg, h = _create(lib, ["-n", "100", "-N", "10", "--no-such-option", "1"])
    assert rc < 0 and not h.value and msg
    rc, msg, h = _create(lib, pansim_amd.config1_args(100, 10, 1, 0, extra=["-p", "does/not/exist.json"]))
    assert rc < 0 and not h.value and "exist" in msg
    rc, msg, h = _create(lib, pansim_amd.config1_args(100, 10, 1, 0, extra=["--quarantinePolicy", "x"]))
    assert rc < 0 and not h.value


@pytest.mark.skipif(gpu_available(), reason="checks the behaviour without a HIP device")
def test_no_device_fails_loudly(lib):
    rc, msg, h = _create(lib, pansim_amd.config1_args(200, 20, 1, 0))
    assert rc == -8 and not h.value, (rc, msg)  # PANSIM_E_DEVICE
    assert msg
    with pytest.raises(pansim_amd.PanSimError):
        pansim_amd.Simulation(pansim_amd.config1_args(200, 20, 1, 0))
