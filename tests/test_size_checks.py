"""The checks of tests/size_checks.py, which the HIP path runs at the full
size, pinned on the oracle (CPU) at a small size: they pass on a correct
run at the same relative checkpoints, and each fails on a planted mistake
(an occupancy or a sum off by one unit, an agent at a location that is not
its own, a statistics entry off by one)."""
import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim
from tests.size_checks import check_locations, check_midnight_stats, check_occupancy_and_sums, tables_from


class Planted:
    """An OracleSim view with one planted mistake in what it reports."""

    def __init__(self, o, what):
        self.o, self.what = o, what

    def __getattr__(self, name):
        return getattr(self.o, name)

    def field(self, name):
        f = self.o.field(name).copy()
        if self.what == "count" and name == "locCount":
            f[int(np.argmax(f))] += 1
        return f

    def location_sums(self, v):
        s = self.o.location_sums(v).copy()
        if self.what == "sum" and v == 0:
            s[int(np.argmax(s))] += 1
        return s

    def day_stats(self, d):
        st = list(self.o.day_stats(d))
        if self.what == "stats":
            st[2] += 1
        return st


@pytest.fixture(scope="module")
def run():
    args = pansim_amd.with_data_dir(pansim_amd.synthetic_args(4000, 1, 2, 3, extra=["-k", "0.01", "--synthInfected", "0.01"]))
    o = OracleSim(args)
    yield o, tables_from(o)
    o.close()


def test_checks_pass_on_the_oracle(run):
    o, t = run
    for target in (144 + 61, 2 * 144 + 1, 2 * 144 + 107):
        o.steps(target - o.timestamp)
        loc, state = check_occupancy_and_sums(o, t)
        check_locations(loc, state, t)
        if target % 144 == 1:
            check_midnight_stats(o, state, t)


def _after_midnight(o):
    if o.timestamp % 144 != 1:
        o.steps(144 - o.timestamp % 144 + 1)


@pytest.mark.parametrize("what,message", [("count", "occupancy"), ("sum", "variant 0 sums")])
def test_sum_checks_catch_planted_mistakes(run, what, message):
    o, t = run
    _after_midnight(o)
    with pytest.raises(AssertionError, match=message):
        check_occupancy_and_sums(Planted(o, what), t)


def test_location_checks_catch_planted_mistakes(run):
    o, t = run
    _after_midnight(o)
    loc, state = check_occupancy_and_sums(o, t)
    begin, count, locs = t["begin"], t["count"], t["locs"]
    assert t["cem"] == o.cemetery  # the cemetery is the last location
    # agent 0 at a location from another agent's list, not its own
    own = set(locs[begin[0]:begin[0] + count[0]].tolist())
    bad = loc.copy()
    bad[0] = next(int(x) for x in locs[count[0]:] if int(x) not in own and int(x) != t["cem"])
    with pytest.raises(AssertionError, match="away from their possible locations"):
        check_locations(bad, state, t)
    # a living agent at the cemetery
    living = int(np.nonzero(t["wb"][state] != 4)[0][0])
    bad = loc.copy()
    bad[living] = t["cem"]
    with pytest.raises(AssertionError, match="living agent at the cemetery"):
        check_locations(bad, state, t)


def test_stats_check_catches_a_planted_mistake(run):
    o, t = run
    _after_midnight(o)
    state = o.field("state").astype(np.int64)
    check_midnight_stats(o, state, t)
    with pytest.raises(AssertionError, match="state counts"):
        check_midnight_stats(Planted(o, "stats"), state, t)
