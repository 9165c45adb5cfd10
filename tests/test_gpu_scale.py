"""The plain GPU scale workload at its size: 100M agents from the synthetic
generator (about 46M locations), 2 weeks, quarantine policy 3, testing on,
uint32 location ids, exact fixed-point exposure sums, about 13 GB of HBM.
The oracle does not run at this size; the size-independent properties of
tests/size_checks.py (pinned on the oracle by tests/test_size_checks.py) are
checked one step after a midnight block, at the morning activity change of a
weekday (08:00, millions of agents re-evaluated) and at the last step."""
import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim
from tests.size_checks import check_locations, check_midnight_stats, check_occupancy_and_sums, tables_from

pytestmark = pytest.mark.gpu

N_AGENTS = 100_000_000


def test_scale_100M_two_weeks():
    sim = pansim_amd.Simulation(pansim_amd.synthetic_args(N_AGENTS, 2, 0, 1))
    try:
        ref = OracleSim(pansim_amd.with_data_dir(pansim_amd.synthetic_args(2000, 2, 0, 1)))
        try:
            t = tables_from(ref, sim)
        finally:
            ref.close()
        assert sim.n_agents == N_AGENTS and sim.n_locations > 40_000_000
        for target in (2 * 144 + 1, 8 * 144 + 48, 14 * 144):
            sim.steps(target - sim.timestamp)
            evaluated, _ = sim.movement_counters(reset=True)
            loc, state = check_occupancy_and_sums(sim, t)
            check_locations(loc, state, t)
            if target % 144 == 1:
                check_midnight_stats(sim, state, t)
            print(f"checked at step {target}: {sim.n_agents} agents, {sim.n_locations} locations", flush=True)
            del loc, state
        assert sim.num_days == 14
    finally:
        sim.close()
