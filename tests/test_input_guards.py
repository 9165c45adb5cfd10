"""Inputs the engine refuses at pansim_create, before any device work (so
these run without a GPU): agent infectiousness values that the exact 2^-32
fixed-point location sums cannot hold (a state value times a variant
multiplier that is non-zero and below 2^-9), and infectious susceptible or
first-infected states (the fused infection pass assumes they carry none)."""
import json
import os
import shutil

import pytest

import pansim_amd

PROG = os.path.join(pansim_amd.DATA_DIR, "inputConfigFiles", "progressions")
E_UNSUPPORTED = -9


def _progression(tmp_path, state, value):
    d = tmp_path / "progressions"
    shutil.copytree(PROG, d)
    cfg = json.load(open(d / "transition_config.json"))
    for s in cfg["states"]:
        if s["stateName"] == state:
            s["infectious"] = value
    (d / "transition_config.json").write_text(json.dumps(cfg))
    return str(d / "transition_config.json")


@pytest.mark.parametrize("state,value", [("I2", 0.001), ("I5_h", 1e-6), ("S", 0.5), ("E", 0.25)])
def test_refused_infectiousness(tmp_path, state, value):
    args = pansim_amd.config1_args(200, 50, 1, 0, extra=["-P", _progression(tmp_path, state, value)])
    with pytest.raises(pansim_amd.PanSimError) as e:
        pansim_amd.Simulation(args)
    assert e.value.code == E_UNSUPPORTED, str(e.value)


def test_smallest_accepted_value_reaches_the_device(tmp_path):
    # 2^-9 times every variant multiplier (>= 1) is representable: the check
    # passes, and creation goes on to the device (absent here, or present on
    # a GPU box where it succeeds)
    args = pansim_amd.config1_args(200, 50, 1, 0, extra=["-P", _progression(tmp_path, "I2", 2.0 ** -9)])
    try:
        pansim_amd.Simulation(args).close()
    except pansim_amd.PanSimError as e:
        assert e.code != E_UNSUPPORTED, str(e)
