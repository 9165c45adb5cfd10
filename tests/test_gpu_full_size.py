"""The size-independent properties of tests/size_checks.py on the HIP path at
the full size of the throughput measurement: the 10M-agent synthetic
household population (4,603,331 locations, all policies on), after a
daytime step of day 1, one step after the second midnight block, at the
evening activity changes of day 2 (17:50), one step after the last midnight block of the four weeks and at their last step.  The oracle
does not run at this size here: a small oracle population built from the
same input files supplies the state table and the variant multipliers, and
the agents' possible-location lists are the ones held on the device."""
import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim
from tests.size_checks import check_locations, check_midnight_stats, check_occupancy_and_sums, tables_from

pytestmark = pytest.mark.gpu

N_AGENTS = 10_000_000


def test_device_location_lists_equal_the_oracles():
    # the lists the full-size check reads from the device are the population's
    args = pansim_amd.synthetic_args(3000, 1, 2, 4)
    sim = pansim_amd.Simulation(args)
    ref = OracleSim(pansim_amd.with_data_dir(args))
    try:
        begin, count, locs = sim.possible_locations()
        off, rlocs, _ = ref.csr()
        assert np.array_equal(begin, off[:-1]) and np.array_equal(count, np.diff(off))
        assert np.array_equal(locs, rlocs)
        assert sim.n_locations - 1 == ref.cemetery
    finally:
        sim.close()
        ref.close()


@pytest.fixture(scope="module")
def full():
    sim = pansim_amd.Simulation(pansim_amd.synthetic_args(N_AGENTS, 4, 0, 1))
    # the state table and variant multipliers come from the same input files
    ref = OracleSim(pansim_amd.with_data_dir(pansim_amd.synthetic_args(2000, 4, 0, 1)))
    try:
        t = tables_from(ref, sim)
    finally:
        ref.close()
    assert sim.n_agents == N_AGENTS and len(t["begin"]) == N_AGENTS
    yield sim, t
    sim.close()


@pytest.mark.parametrize("target", [144 + 61, 2 * 144 + 1, 2 * 144 + 107, 27 * 144 + 1, 28 * 144 - 1])
def test_full_size_invariants(full, target):
    sim, t = full
    assert sim.timestamp < target
    sim.steps(target - sim.timestamp)
    loc, state = check_occupancy_and_sums(sim, t)
    check_locations(loc, state, t)
    if target % 144 == 1:
        check_midnight_stats(sim, state, t)
