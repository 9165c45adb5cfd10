"""The closure actions that act on movement, pinned by hand on schedules
written here (tests/mini_world.py):

* HolidayMode (closurePolicies.h:273-285, movementPolicies.h:1714): while the
  rule is active every agent follows the Sunday schedule;
* ReduceMovement (closurePolicies.h:392-409, agentType.cpp:28-109):
  setStayHome(p) walks each schedule slot window by window (a window = the
  run of events with the start time of its first one; the run is not stopped
  at the slot's end) and raises a home event's chance to c + (1 - c) p, or
  inserts a home event of chance p in front of the window, then multiplies
  the other chances by (1 - p); unsetStayHome(p) divides them by (1 - p)
  and turns the home chance back into (c - p) / (1 - p), erasing it when it
  reaches 0.  The expected tables below are worked out by hand, including
  the windows that run on into the next weekday's slot.
The rules are "afterDays" rules: active on days threshold .. threshold +
openAfter - 1 (closurePolicies.h:117-125), switched at the midnight block.
"""
from __future__ import annotations

import json
import os

import numpy as np

from oracle.oracle import OracleSim
from tests.mini_world import HOME, WORK, MiniWorld, agent_type, event, person, place, schedule

SHOP7, SHOP8 = 7, 8


def _rules(world, name, threshold, open_after, parameter="0"):
    path = os.path.join(world.dir, "rules.json")
    with open(path, "w") as fh:
        json.dump({"rules": [{"name": name, "conditionType": "afterDays", "threshold": threshold,
                              "threshold2": threshold, "openAfter": open_after, "closeAfter": -1,
                              "parameter": parameter, "locationTypesToClose": []}]}, fh)
    return ["--enableClosures", "1", "--closures", path]


def _places():
    return [place("H0", HOME), place("W0", WORK), place("S7", SHOP7), place("S8", SHOP8)]


LOCS = [(HOME, "H0"), (WORK, "W0"), (SHOP7, "S7"), (SHOP8, "S8")]


def test_holiday_mode_follows_the_sunday_schedule():
    """WEEKDAYS: work 8:00-8:10 for 2h; SUNDAY: shop8 10:00-10:10 for 1h.  The
    rule is active on day 1 only: on Monday (day 0) and Wednesday (day 2) the
    agent is at work at 8:00; on Tuesday (day 1) it is home at 8:00 and at
    shop8 from 10:00 (Sunday's schedule)."""
    t = agent_type(1, [schedule("W", "WEEKDAYS", [event(WORK, 1, 8, 8.1, 2)]),
                       schedule("W", "SUNDAY", [event(SHOP8, 1, 10, 10.1, 1)])])
    w = MiniWorld(_places(), [person(1, LOCS)], [t])
    o = OracleSim(w.args(extra=_rules(w, "HolidayMode", 1, 1)))
    at = {}
    try:
        for day in range(3):
            for s in (48, 60):
                o.steps(day * 144 + s + 1 - o.timestamp)
                at[(day, s)] = int(o.field("location")[0])
    finally:
        o.close()
    L = w.loc
    assert at == {(0, 48): L("W0"), (0, 60): L("W0"), (1, 48): L("H0"), (1, 60): L("S8"),
                  (2, 48): L("W0"), (2, 60): L("W0")}


def _table(o):
    off, ty, ch, tm = o.schedule()
    return off.tolist(), [(int(ty[k]), float(ch[k]), int(tm[k][0]), int(tm[k][1])) for k in range(len(ty))]


def test_reduce_movement_schedule_tables():
    """Type 1 (every day): window 8:00 {shop7 .5, shop8 .5}, window 12:00
    {work 1}.  Type 2 (weekdays): one window, shop8 8:00-17:00, chance 1.
    p = 0.5, active on days 1-2.

    setStayHome, type 1, each day: 8:00 gets home .5 in front, the shops
    .25 each; 12:00 gets home .5 in front, work .5 (the 12:00 window ends
    before the next day's 8:00 events, different start).
    setStayHome, type 2: Monday's window runs on through Friday's event (all
    start at 8:00): one home .5 is inserted in front of Monday and all five
    shop8 events are halved; then Tuesday's window (Tuesday..Friday) gets its
    own home and is halved again, and so on: shop8 = .5, .25, .125, .0625,
    .03125 from Monday to Friday, each day led by a home of .5.
    unsetStayHome: type 2 gets its original table back (every window runs to
    Friday, the shops are divided by .5 as often as they were multiplied and
    each home, (.5 - .5)/.5 = 0, is erased).  Type 1: in the 8:00 window the
    home is erased and the shops go back to .5; the next window then starts
    at the old window end, which after the erase (agentType.cpp:45-53 does
    not move it back) is the work event: the 12:00 home is skipped and stays
    at .5, while work is divided by .5 to 1.  Each day of type 1 keeps four
    events: shop7 .5, shop8 .5, home .5 and work 1 (12:00)."""
    t1 = agent_type(1, [schedule("W", "ALL", [event(SHOP7, .5, 8, 8.1, 1), event(SHOP8, .5, 8, 8.1, 1),
                                             event(WORK, 1, 12, 12.1, 1)])])
    t2 = agent_type(2, [schedule("W", "WEEKDAYS", [event(SHOP8, 1, 8, 17, 1)])])
    w = MiniWorld(_places(), [person(1, LOCS), person(2, LOCS)], [t1, t2])
    o = OracleSim(w.args(extra=_rules(w, "ReduceMovement", 1, 2, "0.5")))
    try:
        before = _table(o)
        o.steps(145)  # through the midnight of day 1: the rule switches on
        during = _table(o)
        o.steps(3 * 144 + 1 - o.timestamp)  # through the midnight of day 3: off again
        after = _table(o)
    finally:
        o.close()
    # type 1: 7 days x 3 events, then type 2: 5 days x 1 event (W slots only)
    d1 = [(SHOP7, .5, 8, 0), (SHOP8, .5, 8, 0), (WORK, 1.0, 12, 0)]
    exp_before_ev = d1 * 7 + [(SHOP8, 1.0, 8, 0)] * 5
    exp_before_off = [3 * min(s, 7) for s in range(29)] + [21 + min(s, 5) for s in range(1, 29)]
    assert before == (exp_before_off, exp_before_ev)
    d1_set = [(HOME, .5, 8, 0), (SHOP7, .25, 8, 0), (SHOP8, .25, 8, 0), (HOME, .5, 12, 0), (WORK, .5, 12, 0)]
    t2_set = []
    for k in range(5):
        t2_set += [(HOME, .5, 8, 0), (SHOP8, .5 ** (k + 1), 8, 0)]
    exp_during_off = [5 * min(s, 7) for s in range(29)] + [35 + 2 * min(s, 5) for s in range(1, 29)]
    assert during == (exp_during_off, d1_set * 7 + t2_set)
    d1_after = [(SHOP7, .5, 8, 0), (SHOP8, .5, 8, 0), (HOME, .5, 12, 0), (WORK, 1.0, 12, 0)]
    exp_after_off = [4 * min(s, 7) for s in range(29)] + [28 + min(s, 5) for s in range(1, 29)]
    assert after == (exp_after_off, d1_after * 7 + [(SHOP8, 1.0, 8, 0)] * 5)


def test_reduce_movement_shares_while_active():
    """Type 1 at 8:00 while the rule is active: home .5, shop7 .25, shop8 .25
    (cumulative pick over the table above); before and after it: shop7 .5,
    shop8 .5, nobody stays home."""
    t1 = agent_type(1, [schedule("W", "ALL", [event(SHOP7, .5, 8, 8.1, 1), event(SHOP8, .5, 8, 8.1, 1),
                                             event(WORK, 1, 12, 12.1, 1)])])
    n = 4000
    w = MiniWorld(_places(), [person(1, LOCS) for _ in range(n)], [t1])
    o = OracleSim(w.args(extra=_rules(w, "ReduceMovement", 1, 1, "0.5")))
    shares = {}
    try:
        for day in range(3):
            o.steps(day * 144 + 49 - o.timestamp)
            loc = o.field("location")
            shares[day] = [float((loc == w.loc(x)).mean()) for x in ("H0", "S7", "S8")]
    finally:
        o.close()
    tol = 4.5 * np.sqrt(0.25 / n)
    for day, exp in ((0, [0, .5, .5]), (1, [.5, .25, .25]), (2, [0, .5, .5])):
        assert np.all(np.abs(np.array(shares[day]) - exp) <= tol), (day, shares[day])
