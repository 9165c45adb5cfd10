"""Helpers comparing the HIP path (pansim_amd.Simulation) with the CPU
oracle (oracle/oracle.py) on identical inputs and seeds."""
from __future__ import annotations

import numpy as np

import pansim_amd
from oracle.oracle import OracleSim

NEVER = 0xFFFFFFFF
EXACT_FIELDS = [f for f in pansim_amd.AGENT_FIELDS if f != "stepsUntilMove"] + pansim_amd.LOCATION_FIELDS


def next_move_steps(sum_or_rel: np.ndarray, ts: int, steps_per_day: int) -> np.ndarray:
    """Steps until the next full movement evaluation, clamped at the next
    midnight (where planLocations resets every countdown)."""
    until_midnight = steps_per_day - (ts % steps_per_day)
    v = sum_or_rel.astype(np.int64)
    return np.minimum(v, until_midnight)


def compare_state(gpu: pansim_amd.Simulation, ora: OracleSim, where: str, fields=None, countdown=True):
    assert gpu.timestamp == ora.timestamp, where
    ts = gpu.timestamp
    mism = []
    for f in fields or EXACT_FIELDS:
        a = gpu.field(f)
        b = ora.field(f)
        if not np.array_equal(a, b):
            idx = np.nonzero(a != b)[0]
            mism.append(f"{f}: {len(idx)} differ, first {idx[:5].tolist()} gpu={a[idx[:5]].tolist()} oracle={b[idx[:5]].tolist()}")
    # countdown: the oracle stores stepsUntilMove, the GPU the absolute wake step
    spd = gpu.steps_per_day
    if countdown and ts % spd != 0:
        g = gpu.field("stepsUntilMove")
        o = ora.field("stepsUntilMove")
        g_next = next_move_steps(np.where(g == NEVER, np.int64(1) << 40, g.astype(np.int64)), ts, spd)
        o_next = next_move_steps(o.astype(np.int64), ts, spd)
        if not np.array_equal(g_next, o_next):
            idx = np.nonzero(g_next != o_next)[0]
            mism.append(f"stepsUntilMove: {len(idx)} differ, first {idx[:5].tolist()} gpu={g_next[idx[:5]].tolist()} oracle={o_next[idx[:5]].tolist()}")
    for v in range(gpu.n_variants):
        a = gpu.location_sums(v)
        b = ora.location_sums(v)
        if not np.array_equal(a, b):
            idx = np.nonzero(a != b)[0]
            mism.append(f"locSum[{v}]: {len(idx)} differ, first {idx[:5].tolist()}")
    assert not mism, f"{where} (ts={ts}):\n" + "\n".join(mism)


def compare_days(gpu, ora):
    assert gpu.num_days == ora.num_days
    for d in range(gpu.num_days):
        a, b = gpu.day_stats(d), ora.day_stats(d)
        assert a == b, f"day {d} statistics differ:\n gpu    {a}\n oracle {b}"


def run_parity(args, checkpoints, flags_per_day=None, countdown=True):
    """Steps both implementations through `checkpoints` (cumulative step
    counts, or ("day", flags) for runForDay) comparing every field
    (countdown=False: not the RealMovement countdown, which the other
    movement policies leave alone)."""
    gpu = pansim_amd.Simulation(args)
    ora = OracleSim(args)
    try:
        compare_state(gpu, ora, "after init", countdown=countdown)
        done = 0
        for cp in checkpoints:
            if isinstance(cp, tuple):
                a = gpu.run_for_day(cp[1])
                b = ora.run_for_day(cp[1])
                assert a == b, f"runForDay {cp[1]} statistics differ:\n gpu    {a}\n oracle {b}"
                compare_state(gpu, ora, f"after runForDay {cp[1]}", countdown=countdown)
                continue
            gpu.steps(cp - done)
            ora.steps(cp - done)
            done = cp
            compare_state(gpu, ora, f"after {cp} steps", countdown=countdown)
        compare_days(gpu, ora)
        return gpu, ora
    except Exception:
        gpu.close()
        ora.close()
        raise
