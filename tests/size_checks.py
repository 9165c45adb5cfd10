"""Properties of a simulation state that hold at any population size, written
against the accessors that the HIP path (pansim_amd.Simulation) and the
oracle (oracle.oracle.OracleSim) share, so that they are pinned on the
oracle at small sizes (test_size_checks.py) and applied to the HIP path at
the full size (test_gpu_full_size.py):

- per-location occupancy equals the histogram of the agents' locations
  (the reference recounts it every step, util.cpp:39-110);
- the exact per-location infectiousness sums equal a recount of the agents
  present, in 2^-32 fixed point, for every variant (util.h:47-103);
- every agent is at one of its own possible locations or at the cemetery
  (movementPolicies.h:69-122, 337-341), and only dead agents are there;
- one step after a midnight block, the first statistics entries are the
  per-state agent counts of that midnight (simulation.h:473-490).
"""
import numpy as np

WB_DEAD = 4


def tables_from(ref, sim=None):
    """State table and variant multipliers from `ref` (an OracleSim built from
    the same input files: they do not depend on the population), and the
    population's possible-location lists from `sim` (the simulation under
    test: the HIP path's lists as held on the device), or from `ref` when
    `sim` is None.  The cemetery is the last location."""
    wb, inf = ref.state_table()
    if sim is None:
        off, locs, _ = ref.csr()
        off = off.astype(np.int64)
        begin, count, nl = off[:-1], np.diff(off), ref.n_locations
    else:
        begin, count, locs = sim.possible_locations()
        begin, count, nl = begin.astype(np.int64), count.astype(np.int64), sim.n_locations
    return dict(begin=begin, count=count, locs=locs, wb=wb, inf=inf, vm=ref.variant_multipliers(), cem=nl - 1, nl=nl)


def check_occupancy_and_sums(sim, t):
    loc = sim.field("location").astype(np.int64)
    nl = t["nl"]
    assert np.array_equal(np.bincount(loc, minlength=nl), sim.field("locCount").astype(np.int64)), "occupancy"
    state = sim.field("state").astype(np.int64)
    var = sim.field("variant").astype(np.int64)
    x = t["inf"][state] * t["vm"][var]  # float32 product, as in the simulation
    q = (x.astype(np.float64) * 4294967296.0).astype(np.int64)
    assert (q > 0).sum() > 0, "no infectious agent: nothing was checked"
    for v in range(len(t["vm"])):
        # partial sums stay below 2^53, so the float64 histogram is exact
        ref = np.bincount(loc, weights=np.where(var == v, q, 0).astype(np.float64), minlength=nl).astype(np.int64)
        assert np.array_equal(ref, sim.location_sums(v)), f"variant {v} sums"
    return loc, state


def check_locations(loc, state, t):
    begin, count, locs = t["begin"], t["count"], t["locs"]
    n = len(loc)
    # the lists are stored one after the other in agent order
    assert begin[0] == 0 and np.array_equal(begin[1:], begin[:-1] + count[:-1]) and len(locs) == begin[-1] + count[-1]
    match = locs == np.repeat(loc.astype(np.uint32), count)
    own = np.zeros(n, dtype=bool)
    own[np.repeat(np.arange(n, dtype=np.int64), count)[match]] = True
    at_cem = loc == t["cem"]
    assert np.all(own | at_cem), f"agents away from their possible locations: {np.nonzero(~(own | at_cem))[0][:10]}"
    assert np.all(t["wb"][state[at_cem]] == WB_DEAD), "living agent at the cemetery"


def check_midnight_stats(sim, state, t, steps_per_day=144):
    """One step after a midnight block: the statistics were taken at its
    start, and since then only that step's infections (susceptible -> first
    infected state) changed any agent's state."""
    ts = sim.timestamp
    assert ts % steps_per_day == 1
    st = sim.day_stats(sim.num_days - 1)
    ns = len(t["wb"])
    counts = np.bincount(state, minlength=ns)
    new = sim.field("infectedTimestamp") == ts - 1
    before = counts.copy()
    if new.any():
        np.subtract.at(before, state[new], 1)
        before[0] += int(new.sum())
    assert st[:ns] == before.tolist(), "state counts of the statistics"
    assert sum(st[:ns]) == len(state)
