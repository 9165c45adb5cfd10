"""The oracle's non-COVID hospitalisation pass (Simulation::otherDisease,
simulation.h:169-423) pinned by the structure its definition fixes, over a
run long enough for stays to start and end:

- a stay starts at a midnight block and lasts a whole number of days, at
  least one (length = max(1, geometric draw) days, :405-413);
- an agent already dead or in a COVID hospital state, or still in a stay,
  is not hospitalised again (:370-374);
- with the sudden-death branch switched off (:378, `&& false`), a non-COVID
  death (state D2) happens only at the midnight where the agent's stay
  ends (:352-365), so its worst-state timestamp is that midnight.
The dates are read from the agents' fields at every midnight."""
import numpy as np

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144
NEVER = 0xFFFFFFFF


def test_stays_and_non_covid_deaths_follow_the_rules():
    args = pansim_amd.synthetic_args(20000, 10, 5, 2, extra=["-k", "0.01", "--synthInfected", "0.01"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        names = o.state_names()
        d2 = names.index("D2")
        wb, _ = o.state_table()
        starts = ends = deaths = 0
        prev_until = o.field("hospitalizedUntilTimestamp").copy()
        prev_wb = wb[o.field("state").astype(np.int64)]
        for day in range(1, 70):
            o.steps(day * SPD + 1 - o.timestamp)
            ts = day * SPD
            hts = o.field("hospitalizedTimestamp")
            until = o.field("hospitalizedUntilTimestamp")
            new = hts == ts
            if new.any():
                starts += int(new.sum())
                length = until[new].astype(np.int64) - ts
                assert np.all(length >= SPD) and np.all(length % SPD == 0), "stays are whole days"
                # not admitted while dead, in a COVID hospital state, or in a stay
                assert np.all(prev_wb[new] != 4) and np.all(prev_wb[new] != 3)
                assert np.all(~((prev_until[new] != NEVER) & (prev_until[new] > ts)))
            state = o.field("state").astype(np.int64)
            died = (state == d2) & (o.field("worstStateTimestamp") == ts) & (o.field("worstState") == d2)
            if died.any():
                deaths += int(died.sum())
                assert np.all(until[died] == ts), "non-COVID deaths only at the end of a stay"
            ends += int((until == ts).sum())
            prev_until = until.copy()
            prev_wb = wb[state]
        # the pass ran often enough to test anything (about 1500 stays and
        # 50 deaths in this run)
        assert starts > 200 and ends > 200 and deaths > 10
    finally:
        o.close()
