"""The oracle's testing pass (DetailedTesting::performTests,
testingPolicies.h:59-281) pinned by what its probabilities imply, with the
testing probabilities chosen so that one term acts at a time:

- "0.5,0,0,0,0,0" (only the random term): every agent that may be tested
  (not dead, not diagnosed, no test in the last testingRepeatDelay days,
  testingPolicies.h:170-180) is tested with probability 0.5, so the day's
  test count is Binomial(eligible, 0.5);
- "0,1,0,0,0,0" (only the home term, probability exactly 1): the tested
  agents are exactly the eligible agents whose home has a resident
  diagnosed within the last day (flagLocations, testingPolicies.h:59-148),
  and the others are not tested at all.
The eligible set and the flagged homes are recounted from the agent fields
before the midnight block; the test count is the day's statistics entry
(simulation.h:473-665) and the agents' lastTest field."""
import json
import os

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144
DELAY = 5  # testingRepeatDelay default (testingPolicies.h:318-335)
WB_DEAD = 4
NEVER = 0xFFFFFFFF


def _home_type():
    lt = json.load(open(os.path.join(pansim_amd.DATA_DIR, "inputConfigFiles", "locationTypes.json")))
    return int(lt["home"])


def _sim(probs):
    args = pansim_amd.synthetic_args(20000, 2, 7, 2, extra=[
        "-k", "0.05", "--synthInfected", "0.05", "--testingProbabilities", probs])
    return OracleSim(pansim_amd.with_data_dir(args))


def _before_midnight(o, day):
    """Fields just before the midnight block that starts `day`."""
    o.steps(day * SPD - o.timestamp)
    ts = o.timestamp  # the midnight block runs at the start of this step
    wb, _ = o.state_table()
    last = o.field("lastTest")
    eligible = ~((ts > DELAY * SPD) & (last != NEVER) & (last > ts - DELAY * SPD))
    eligible &= (wb[o.field("state").astype(np.int64)] != WB_DEAD) & (o.field("diagnosed") == 0)
    return ts, eligible


def _tests_of_the_day(o, ts):
    o.steps(1)
    st = o.day_stats(o.num_days - 1)
    ns = len(o.state_names())
    tested = o.field("lastTest") == ts
    assert st[ns + 1] == int(tested.sum())  # the statistics entry counts the same agents
    return tested


@pytest.mark.parametrize("day", [2, 4, 7])
def test_random_term_is_binomial_over_the_eligible(day):
    o = _sim("0.5,0,0,0,0,0")
    try:
        ts, eligible = _before_midnight(o, day)
        tested = _tests_of_the_day(o, ts)
        assert not np.any(tested & ~eligible), "an ineligible agent was tested"
        e = int(eligible.sum())
        t = int(tested.sum())
        assert e > 1000
        assert abs(t - 0.5 * e) <= 5.0 * np.sqrt(0.25 * e), (t, e)
    finally:
        o.close()


def test_home_term_tests_exactly_the_flagged_households():
    o = _sim("0,1,0,0,0,0")
    try:
        off, locs, types = o.csr()
        n = o.n_agents
        home_t = _home_type()
        # each agent's home: the first home-type entry of its list
        home = np.full(n, -1, dtype=np.int64)
        owner = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
        is_home = types == home_t
        first = np.unique(owner[is_home], return_index=True)
        home[first[0]] = locs[is_home][first[1]]
        checked = 0
        for day in (3, 5, 8):
            ts, eligible = _before_midnight(o, day)
            diag = o.field("diagnosedTimestamp")
            recent = (diag != NEVER) & (diag > ts - SPD)
            flagged = set(home[recent & (home >= 0)].tolist())
            expect = eligible & np.isin(home, list(flagged)) & (home >= 0)
            tested = _tests_of_the_day(o, ts)
            assert np.array_equal(tested, expect), (day, int(tested.sum()), int(expect.sum()))
            checked += int(expect.sum())
        assert checked > 50  # some households were flagged
    finally:
        o.close()
