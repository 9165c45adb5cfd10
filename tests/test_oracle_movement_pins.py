"""The oracle's movement routine (RealMovementOps::doMovement,
movementPolicies.h:258-1041) pinned by hand-computed trajectories of agents on
schedules written here (tests/mini_world.py).  Every expected location and
countdown below is worked out from the reference's rules by hand, in the
comments next to it; none is read back from the oracle.

Notation: step s of a day is the time s*10 min (dt = 10); "next evaluation at
step t" means the agent's countdown reaches 0 there.  A full evaluation that
sets stepsUntilMove = X and falls through to the final decrement (:1040)
evaluates again X steps later; one that returns early (quarantine, hospital)
X + 1 steps later.  At every midnight planLocations sets stepsUntilMove = 0
(curfewEnd, :1446), so every agent is evaluated at 00:00.

Rules exercised (reference lines):
  * schedule window scan (:585-604): an event is active for start <= t < end;
    the first later event ends the scan (activeEnd); when the last event of
    the slot is active, activeEnd = eventsEnd and the stay rule reads the
    event one past the slot -- the first event of the next slot in offset
    order (agentType.cpp:13-15, 20-26);
  * Case 1 (:634-656): no active and no later event -> home until midnight;
  * Cases 2/4 (:658-867): a single active event -> its type for its duration,
    then the three stay rules against the next event (:793-815): arrival
    before its start -> the duration; arrival after its end ->
    (end - now) - 1 steps; otherwise the duration plus a uniform extra below
    the steps from arrival to its end; several active events -> one picked by
    cumulative chance; a closed destination -> the conditional re-pick among
    the others, then home; work is never closed (:708-709);
  * Case 3 (:870-920): nothing active but a later event -> wait
    (start - now) + uniform(length) steps: < 30 min stay, < 60 min public
    space, else home;
  * dead -> cemetery for good (:336-341); WB S -> hospital until midnight,
    diagnosed on arrival (:405-443).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import OracleSim
from tests.mini_world import (DOCTOR, HOME, HOSPITAL, PUBLIC, WORK, MiniWorld, agent_type, event, person, place,
                              runs, schedule, trajectory)

UMAX = 0xFFFFFFFF
SHOP7, SHOP8, PUB5 = 7, 8, 5


def _world_places():
    return [place("H0", HOME), place("W0", WORK), place("P0", PUBLIC), place("S7", SHOP7), place("S8", SHOP8),
            place("HOSP", HOSPITAL), place("DOC", DOCTOR), place("R5", PUB5)]


ALL_LOCS = [(HOME, "H0"), (WORK, "W0"), (PUBLIC, "P0"), (SHOP7, "S7"), (SHOP8, "S8"), (HOSPITAL, "HOSP"),
            (DOCTOR, "DOC"), (PUB5, "R5")]


def _run(world, n_steps, **kw):
    o = OracleSim(world.args(**kw))
    try:
        return trajectory(o, n_steps, ("location", "stepsUntilMove", "stayedHome", "diagnosed", "quarantined"))
    finally:
        o.close()


def test_one_day_every_case3_branch_and_the_last_event():
    """Events (type, start-end, duration): work 8:00-8:10 2h; shop7 10:40-10:50
    30 min; shop8 11:30-11:40 1h; shop7 14:00-14:10 30 min.  Every window is
    one step long, so uniform(length) = uniform(1) = 0 and the day is fixed:
      00:00 Case 3 to 8:00: wait 48 >= 6 -> home, next at step 48.
      08:00 work 2h, arrival 10:00 < 10:40 -> 12 steps, next at 60.
      10:00 Case 3 to 10:40: 4 steps, 3 <= 4 < 6 -> public space, next at 64.
      10:40 shop7 30 min, arrival 11:10 < 11:30 -> 3 steps, next at 67.
      11:10 Case 3 to 11:30: 2 < 3 -> stays at shop7, next at 69.
      11:30 shop8 1h, arrival 12:30 < 14:00 -> 6 steps, next at 75.
      12:30 Case 3 to 14:00: 9 >= 6 -> home, next at 84.
      14:00 shop7, the slot's last event: the rule reads the next day's first
            event (work 8:00-8:10); arrival 14:30 is after its end, so
            (8:10 - 14:00) - 1 steps, where the unsigned minute difference
            wraps (2^32 - 350 minutes = 62 h 26 min in the TimeDay's
            unsigned-char hours, 3746 min, 374 steps): 373, after the final
            decrement 372 -- the agent stays until midnight.
      00:00 (next day) home again."""
    t = agent_type(1, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 2), event(SHOP7, 1, 10.4, 10.5, 0.3),
                                             event(SHOP8, 1, 11.3, 11.4, 1), event(SHOP7, 1, 14, 14.1, 0.3)])])
    w = MiniWorld(_world_places(), [person(1, ALL_LOCS)], [t])
    tr = _run(w, 146)
    L = w.loc
    exp = runs([(L("H0"), 0, 47), (L("W0"), 48, 59), (L("P0"), 60, 63), (L("S7"), 64, 68), (L("S8"), 69, 74),
                (L("H0"), 75, 83), (L("S7"), 84, 143), (L("H0"), 144, 145)])
    assert tr["location"][:, 0].tolist() == exp.tolist()
    cd = tr["stepsUntilMove"][:, 0]
    # countdown after the evaluation steps (value set minus the final decrement)
    assert [int(cd[s]) for s in (0, 48, 60, 64, 67, 69, 75, 84, 144)] == [47, 11, 3, 2, 1, 5, 8, 372, 47]
    # 12 steps at work (> 6): not stayed home (:1035-1038); reset at midnight
    assert tr["stayedHome"][143, 0] == 0 and tr["stayedHome"][144, 0] == 1


def test_stay_rules_against_the_next_window():
    """work 8:00-8:10 for 5h: arrival 13:00 is after the next event's end
    (12:10) -> (12:10 - 8:00) - 1 = 24 steps, next at step 72.
    shop7 12:00-12:10 for 1h: arrival 13:00 inside the next window
    [13:00, 13:10] -> 6 + uniform(1) = 6 steps, next at 78.
    shop8 13:00-13:10 for 20 min: the last event; the next slot's first event
    is work 8:00-8:10, arrival 13:20 after its end -> stays until midnight.
    A second agent type has the next window 13:00-14:00: its stay after shop7
    is 6 + uniform(6) steps, so its next evaluation falls on each of steps
    78..83 and on no other."""
    t1 = agent_type(1, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 5), event(SHOP7, 1, 12, 12.1, 1),
                                             event(SHOP8, 1, 13, 13.1, 0.2)])])
    t2 = agent_type(2, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 5), event(SHOP7, 1, 12, 12.1, 1),
                                             event(SHOP8, 1, 13, 14, 0.2)])])
    people = [person(1, ALL_LOCS)] + [person(2, ALL_LOCS) for _ in range(240)]
    w = MiniWorld(_world_places(), people, [t1, t2])
    tr = _run(w, 144)
    L = w.loc
    exp = runs([(L("H0"), 0, 47), (L("W0"), 48, 71), (L("S7"), 72, 77), (L("S8"), 78, 143)])
    assert tr["location"][:, 0].tolist() == exp.tolist()
    loc = tr["location"][:, 1:]
    assert (loc[48:72] == L("W0")).all() and (loc[72:78] == L("S7")).all()
    first_s8 = np.argmax(loc == L("S8"), axis=0)
    assert set(first_s8.tolist()) == set(range(78, 84))
    assert min(np.bincount(first_s8 - 78, minlength=6)) >= 20  # each of the 6 delays, 240 agents


def test_weekday_and_weekend_slots_and_case1():
    """WEEKDAYS: work 8:00-8:10 for 2h.  WEEKENDS: shop8 10:00-10:10 for 1h.
    N (another WB state, never used here): shop7 9:00-9:10 -- it only fills
    the slot after Sunday's in offset order.  Day 0 is a Monday.
      Mon-Thu: home; 8:00 work, the last event of the slot; the next slot
        (the next weekday) starts with work 8:00-8:10 and arrival 10:00 is
        after its end -> (8:10 - 8:00) - 1 = 0, and the final decrement wraps
        the countdown: the agent stays at work until midnight.
      Fri: the next slot is Saturday's (shop8 10:00-10:10); arrival 10:00 is
        inside [10:00, 10:10] -> 12 + uniform(1) = 12 steps, next at step 60;
        10:00 nothing active and nothing later -> Case 1, home until
        midnight.
      Sat: 00:00 Case 3 to 10:00 (60 steps) -> home; 10:00 shop8; the next
        slot is Sunday's (10:00-10:10), arrival 11:00 after its end -> 0 ->
        wraps: stays until midnight.
      Sun: as Saturday; the next slot is the N slot of Monday (9:00-9:10):
        arrival 11:00 after its end -> the wrapped difference, 403 steps:
        stays until midnight."""
    t = agent_type(1, [schedule("W", "WEEKDAYS", [event(WORK, 1, 8, 8.1, 2)]),
                       schedule("W", "WEEKENDS", [event(SHOP8, 1, 10, 10.1, 1)]),
                       schedule("N", "ALL", [event(SHOP7, 1, 9, 9.1, 1)])])
    w = MiniWorld(_world_places(), [person(1, ALL_LOCS)], [t])
    tr = _run(w, 7 * 144)
    L = w.loc
    day = {d: tr["location"][d * 144:(d + 1) * 144, 0].tolist() for d in range(7)}
    workday = runs([(L("H0"), 0, 47), (L("W0"), 48, 143)]).tolist()
    for d in range(4):
        assert day[d] == workday, d
    assert day[4] == runs([(L("H0"), 0, 47), (L("W0"), 48, 59), (L("H0"), 60, 143)]).tolist()
    weekend = runs([(L("H0"), 0, 59), (L("S8"), 60, 143)]).tolist()
    assert day[5] == weekend and day[6] == weekend
    cd = tr["stepsUntilMove"][:, 0]
    # Case 1 at 10:00: steps until midnight (1440 - 600) / 10 = 84, minus the final decrement
    assert cd[48] == UMAX and cd[4 * 144 + 48] == 11 and cd[4 * 144 + 60] == 83
    assert cd[6 * 144 + 60] == 402  # 403 set, minus the final decrement
    # startDay 5: day 0 is a Saturday
    tr5 = _run(w, 144, start_day=5)
    assert tr5["location"][:, 0].tolist() == weekend


def _shares(loc_row, ids):
    return np.array([(loc_row == i).mean() for i in ids])


@pytest.mark.parametrize("closed", [(), ("S7",), ("S7", "S8"), ("W0",)])
def test_event_choice_by_cumulative_chance_and_closed_locations(closed):
    """Three events in one window, 8:00-8:10 (shop7 p = .5, shop8 p = .3,
    restaurant5 p = .2, 1h each): the pick is the first event whose cumulative
    chance reaches the uniform draw, so the shares are .5/.3/.2.
    Closed shop7: the agents that picked it draw again with the other chances
    divided by (1 - .5): shop8 .6, restaurant .4 -> overall shop8
    .3 + .5*.6 = .6, restaurant .2 + .5*.4 = .4.
    Closed shop7 and shop8: shop7 pickers re-pick shop8 (.6, closed again ->
    home) or the restaurant (.4); shop8 pickers (picked index 1) re-pick with
    threshold .5/.7 for shop7 (closed -> home), else the restaurant ->
    home .5*.6 + .3*5/7, restaurant .2 + .5*.4 + .3*2/7.
    A closed workplace does not count as closed (work type): shares
    unchanged when W0 is the closed one (it is not in this window)."""
    t = agent_type(1, [schedule("W", "ALL", [event(SHOP7, .5, 8, 8.1, 1), event(SHOP8, .3, 8, 8.1, 1),
                                             event(PUB5, .2, 8, 8.1, 1), event(WORK, 1, 12, 12.1, 1)])])
    places = [place(p["ID"], p["type"], state="OFF" if p["ID"] in closed else "ON") for p in _world_places()]
    n = 6000
    w = MiniWorld(places, [person(1, ALL_LOCS) for _ in range(n)], [t])
    tr = _run(w, 49)
    row = tr["location"][48]
    ids = [w.loc("S7"), w.loc("S8"), w.loc("R5"), w.loc("H0")]
    got = _shares(row, ids)
    exp = {(): [.5, .3, .2, 0], ("W0",): [.5, .3, .2, 0], ("S7",): [0, .6, .4, 0],
           ("S7", "S8"): [0, 0, .2 + .5 * .4 + .3 * 2 / 7, .5 * .6 + .3 * 5 / 7]}[closed]
    tol = 4.5 * np.sqrt(0.25 / n)
    assert np.all(np.abs(got - np.array(exp)) <= tol), (closed, got.tolist(), exp)
    assert np.isclose(got.sum(), 1.0)


def test_single_event_closed_goes_home_but_work_stays():
    """One event at 8:00: shop8 (closed) -> home (:761-773); work (closed) ->
    still work (:708-709)."""
    t1 = agent_type(1, [schedule("W", "ALL", [event(SHOP8, 1, 8, 8.1, 1), event(SHOP7, 1, 12, 12.1, 1)])])
    t2 = agent_type(2, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 1), event(SHOP7, 1, 12, 12.1, 1)])])
    places = [place(p["ID"], p["type"], state="OFF" if p["ID"] in ("S8", "W0") else "ON") for p in _world_places()]
    w = MiniWorld(places, [person(1, ALL_LOCS), person(2, ALL_LOCS)], [t1, t2])
    tr = _run(w, 50)
    assert tr["location"][48].tolist() == [w.loc("H0"), w.loc("W0")]
    # the 1h stay is counted from the original pick in both cases: next at 54
    assert tr["stepsUntilMove"][48].tolist() == [5, 5]


def test_case3_random_delay_spread():
    """Nothing active at 00:00, next event shop7 8:00-9:00: wait
    48 + uniform(6) steps -> the first move falls on each of steps 48..53."""
    t = agent_type(1, [schedule("W", "ALL", [event(SHOP7, 1, 8, 9, 0.3), event(WORK, 1, 12, 12.1, 1)])])
    n = 1200
    w = MiniWorld(_world_places(), [person(1, ALL_LOCS) for _ in range(n)], [t])
    tr = _run(w, 60)
    first = np.argmax(tr["location"] == w.loc("S7"), axis=0)
    assert set(first.tolist()) == set(range(48, 54))
    counts = np.bincount(first - 48, minlength=6)
    assert np.all(np.abs(counts - n / 6) <= 4.5 * np.sqrt(n / 6))


def test_dead_agent_to_cemetery_and_hospital_state():
    """A D1 agent (WB D) goes to the cemetery at its first evaluation with
    stepsUntilMove = UINT_MAX and never returns (:336-341).  An I5_h agent
    (WB S) goes to hospital with stepsUntilMove = steps until midnight = 144
    and no decrement (:405-443); it is infectious (0.1), so it is diagnosed
    there, with diagnosedTimestamp 0 + 1 because it is midnight (:428-430).
    An R_h agent (WB S, not infectious) goes to hospital and is not
    diagnosed."""
    t = agent_type(1, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 1), event(SHOP7, 1, 12, 12.1, 1)])])
    people = [person(1, ALL_LOCS, state="D1"), person(1, ALL_LOCS, state="I5_h"), person(1, ALL_LOCS, state="R_h")]
    w = MiniWorld(_world_places(), people, [t])
    o = OracleSim(w.args())
    try:
        o.steps(1)
        assert o.field("location").tolist() == [w.cemetery, w.loc("HOSP"), w.loc("HOSP")]
        assert o.field("stepsUntilMove").tolist() == [UMAX, 144, 144]
        assert o.field("diagnosed").tolist() == [0, 1, 0]
        assert o.field("diagnosedTimestamp").tolist() == [0, 1, 0]
        o.steps(100)
        assert o.field("location").tolist()[0] == w.cemetery
    finally:
        o.close()


# ---------------------------------------------------------------- quarantine
Q_LEN = 10 * 144  # --quarantineLength 10 days in steps


def _quarantine_world():
    """B (index 0) and C (index 1) are healthy workers; A (index 2, I1,
    infectious) shares B's home H0 and C's workplace W0 and goes to hospital
    at 8:00.  Workers: work 8:00-8:10 for 2h, shop7 12:00-12:10 and 14:00-14:10
    for 30 min.  A: hospital 8:00-8:10 for 1h, then the same shops."""
    shops = [event(SHOP7, 1, 12, 12.1, 0.3), event(SHOP7, 1, 14, 14.1, 0.3)]
    t1 = agent_type(1, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 2)] + shops)])
    t2 = agent_type(2, [schedule("W", "ALL", [event(HOSPITAL, 1, 8, 8.1, 1)] + shops)])
    places = [place("H0", HOME), place("H1", HOME), place("W0", WORK), place("W1", WORK), place("S7", SHOP7),
              place("HOSP", HOSPITAL), place("P0", PUBLIC)]
    common = [(SHOP7, "S7"), (HOSPITAL, "HOSP"), (PUBLIC, "P0")]
    people = [person(1, [(HOME, "H0"), (WORK, "W1")] + common),
              person(1, [(HOME, "H1"), (WORK, "W0")] + common),
              person(2, [(HOME, "H0"), (WORK, "W0")] + common, state="I1")]
    return MiniWorld(places, people, [t1, t2])


@pytest.mark.parametrize("qp", [0, 1, 2, 3])
def test_diagnosis_and_location_quarantine(qp):
    """Worked by hand from :171-245, :446-583, :923-1033 for one day:
    A at 8:00: hospital for 1h (arrival 9:00 < 12:00 -> 6 steps); it is
      infectious and at a hospital, so it is diagnosed (diagnosedTimestamp 48)
      and quarantineAgent(48 + 1440) runs: with policy >= 1 it is quarantined
      (until 1488, 10 days); policy >= 2 also quarantines its home H0 until
      1488; policy >= 3 also its workplace W0 (its previous quarantine end,
      0, is before now).  At 9:00 (step 54) a quarantined agent goes home
      until midnight; with policy 0 it is never quarantined and 9:00 is Case 3
      to 12:00 (18 steps) -> home too, then its shops.
    C at 10:00 (step 60, at W0): with policy >= 3 W0 is under quarantine
      (work type), C is quarantined until 1488 (qTs 60, (1488-60)//144 = 9
      days) and sent home for min(1488 - 60 - 1, 84) = 84 steps; otherwise
      Case 3 to 12:00 -> home, 12:00 shop7 (3 steps), 12:30 Case 3 to 14:00
      (9 steps) -> home, 14:00 shop7 until midnight (last event).
    B at 10:00 (step 60, at W1): Case 3 to 12:00 sends it home; with policy
      >= 2 it has just arrived at a home under quarantine (:923-992), so it is
      quarantined until 1488 (qTs 60, 1428//144 = 9 days) and stays home for
      min(1427, 84) steps; otherwise the shops as C."""
    w = _quarantine_world()
    tr = _run(w, 144, quarantine_policy=qp)
    L = w.loc
    shops_day = lambda h, wk: runs([(L(h), 0, 47), (L(wk), 48, 59), (L(h), 60, 71), (L("S7"), 72, 74),
                                    (L(h), 75, 83), (L("S7"), 84, 143)]).tolist()
    loc = tr["location"]
    if qp >= 2:
        assert loc[:, 0].tolist() == runs([(L("H0"), 0, 47), (L("W1"), 48, 59), (L("H0"), 60, 143)]).tolist()
    else:
        assert loc[:, 0].tolist() == shops_day("H0", "W1")
    if qp >= 3:
        assert loc[:, 1].tolist() == runs([(L("H1"), 0, 47), (L("W0"), 48, 59), (L("H1"), 60, 143)]).tolist()
    else:
        assert loc[:, 1].tolist() == shops_day("H1", "W0")
    if qp >= 1:
        assert loc[:, 2].tolist() == runs([(L("H0"), 0, 47), (L("HOSP"), 48, 53), (L("H0"), 54, 143)]).tolist()
    else:
        assert loc[:, 2].tolist() == runs([(L("H0"), 0, 47), (L("HOSP"), 48, 53), (L("H0"), 54, 71),
                                           (L("S7"), 72, 74), (L("H0"), 75, 83), (L("S7"), 84, 143)]).tolist()
    o = OracleSim(w.args(quarantine_policy=qp))
    try:
        o.steps(144)
        f = {k: o.field(k).tolist() for k in ("diagnosed", "diagnosedTimestamp", "quarantined",
                                               "quarantinedTimestamp", "quarantinedUntilTimestamp",
                                               "daysInQuarantine")}
        lq = o.field("locQuarantineUntil").tolist()
    finally:
        o.close()
    assert f["diagnosed"] == [0, 0, 1] and f["diagnosedTimestamp"] == [0, 0, 48]
    qB, qC, qA = qp >= 2, qp >= 3, qp >= 1
    assert f["quarantined"] == [int(qB), int(qC), int(qA)]
    assert f["quarantinedTimestamp"] == [60 if qB else 0, 60 if qC else 0, 48 if qA else 0]
    assert f["quarantinedUntilTimestamp"] == [48 + Q_LEN if x else 0 for x in (qB, qC, qA)]
    assert f["daysInQuarantine"] == [9 if qB else 0, 9 if qC else 0, 10 if qA else 0]
    exp_lq = [0] * w.cemetery
    if qp >= 2:
        exp_lq[L("H0")] = 48 + Q_LEN
    if qp >= 3:
        exp_lq[L("W0")] = 48 + Q_LEN
    assert lq[:w.cemetery] == exp_lq


def _school_world(g2_state="I1"):
    """One school SC0 with classrooms C0, C1, C2 (ids class_school).  Teachers
    G1 (C0) and G2 (C1), infectious (I1), go to hospital at 8:00 and 9:00; the
    pupil E (age 10, C2) goes to school 8:00-8:10 for 6h; its parents F
    (age 40) and F2 (age 30) work at W2 8:00-8:10 for 2h.  Day 0 is a Monday.
    People order: E, F, F2, G1, G2."""
    shops = [event(SHOP7, 1, 15, 15.1, 0.3), event(SHOP7, 1, 16, 16.1, 0.3)]
    t_work = agent_type(1, [schedule("W", "ALL", [event(WORK, 1, 8, 8.1, 2)] + shops)])
    t_pupil = agent_type(2, [schedule("W", "ALL", [event(3, 1, 8, 8.1, 6)] + shops)])
    t_g1 = agent_type(3, [schedule("W", "ALL", [event(HOSPITAL, 1, 8, 8.1, 1)] + shops)])
    t_g2 = agent_type(4, [schedule("W", "ALL", [event(HOSPITAL, 1, 9, 9.1, 1)] + shops)])
    places = [place("SC0", 3), place("C0_SC0", 33), place("C1_SC0", 33), place("C2_SC0", 33), place("H3", HOME),
              place("H4", HOME), place("H5", HOME), place("W2", WORK), place("S7", SHOP7), place("HOSP", HOSPITAL)]
    common = [(SHOP7, "S7"), (HOSPITAL, "HOSP")]
    people = [person(2, [(HOME, "H3"), (3, "SC0"), (33, "C2_SC0")] + common, age=10),
              person(1, [(HOME, "H3"), (WORK, "W2")] + common, age=40),
              person(1, [(HOME, "H3"), (WORK, "W2")] + common, age=30),
              person(3, [(HOME, "H4"), (33, "C0_SC0")] + common, state="I1"),
              person(4, [(HOME, "H5"), (33, "C1_SC0")] + common, state=g2_state)]
    return MiniWorld(places, people, [t_work, t_pupil, t_g1, t_g2])


@pytest.mark.parametrize("qp", [3, 4])
def test_planning_school_work_quarantine_and_child_care(qp):
    """Worked by hand from planLocations (:1428-1622) and its helpers
    (:1051-1310).  Day 0: G1 is diagnosed at 8:00 (until 48 + 1440 = 1488) and
    quarantines its home and its classroom C0 (policy 3: classroom or work
    whose quarantine and closure have ended); G2 at 9:00 (until 1494) its home
    and C1.  Midnight of day 1 (ts 144):
      policy 4, checkSchoolQuarantine: two classrooms of SC0 are under
        quarantine -> SC0 and all its classrooms until the later end, 1494;
      checkSchoolWorkQuarantine (policy >= 3): E (not quarantined) reads its
        school first: under policy 4 SC0 is quarantined -> E quarantined until
        1494 ((1494 - 144)//144 = 9 days); under policy 3 neither SC0 nor C2 is;
      school/classroom quarantine becomes closure: closedUntil = the
        quarantine end, quarantineUntil = 0 (C0 1488, C1 1494, and SC0, C2
        1494 under policy 4);
      checkUnderageAtHome: E (age <= 14) quarantined -> its home needs care;
      setNoWorkToday: the first adult (27..64) of that home, F, stays home
        (noWorkAgent 1), F2 does not.
    Day 1: F's 8:00 work event becomes home and counts a day in quarantine
    (it is not quarantined); E is home all day under policy 4 and at school
    from 8:00 under policy 3 (SC0 and its own classroom C2 open)."""
    w = _school_world()
    o = OracleSim(w.args(quarantine_policy=qp))
    L = w.loc
    try:
        o.steps(145)  # through day 1's first step (00:00)
        lq = o.field("locQuarantineUntil").tolist()
        cu = o.field("locClosedUntil").tolist()
        f = {k: o.field(k).tolist() for k in ("quarantined", "quarantinedTimestamp", "quarantinedUntilTimestamp",
                                               "daysInQuarantine", "noWorkAgent")}
        o.steps(144 + 51 - 145)  # to 8:30 of day 1
        loc = o.field("location").tolist()
        days = o.field("daysInQuarantine").tolist()
    finally:
        o.close()
    if qp == 4:
        assert [cu[L(x)] for x in ("SC0", "C0_SC0", "C1_SC0", "C2_SC0")] == [1494] * 4
    else:
        assert [cu[L(x)] for x in ("SC0", "C0_SC0", "C1_SC0", "C2_SC0")] == [0, 1488, 1494, 0]
    assert [lq[L(x)] for x in ("SC0", "C0_SC0", "C1_SC0", "C2_SC0")] == [0] * 4
    assert [lq[L("H4")], lq[L("H5")], lq[L("H3")]] == [1488, 1494, 0]
    e_q = qp == 4
    assert f["quarantined"] == [int(e_q), 0, 0, 1, 1]
    assert f["quarantinedTimestamp"][0] == (144 if e_q else 0)
    assert f["quarantinedUntilTimestamp"] == [1494 if e_q else 0, 0, 0, 1488, 1494]
    assert f["daysInQuarantine"] == [9 if e_q else 0, 0, 0, 10, 10]
    assert f["noWorkAgent"] == [0, int(e_q), 0, 0, 0]
    assert loc[0] == (L("H3") if e_q else L("SC0"))
    assert loc[1] == (L("H3") if e_q else L("W2")) and loc[2] == L("W2")
    assert days[1] == (1 if e_q else 0)


def test_one_quarantined_classroom_leaves_the_school_open():
    """As above under policy 4, but G2 is healthy: only C0 is under quarantine
    at the midnight, and checkSchoolQuarantine needs more than one classroom
    (:1297): SC0 stays open, C0 alone becomes closed until 1488, E is not
    quarantined and goes to school, F goes to work."""
    w = _school_world(g2_state="S")
    o = OracleSim(w.args(quarantine_policy=4))
    L = w.loc
    try:
        o.steps(144 + 51)
        cu = o.field("locClosedUntil").tolist()
        q = o.field("quarantined").tolist()
        loc = o.field("location").tolist()
    finally:
        o.close()
    assert [cu[L(x)] for x in ("SC0", "C0_SC0", "C1_SC0", "C2_SC0")] == [0, 1488, 0, 0]
    assert q == [0, 0, 0, 1, 0]
    assert loc[:3] == [L("SC0"), L("W2"), L("W2")]


# ---------------------------------------------------------------- planted mistakes
@pytest.mark.parametrize("code,check", [
    (1, lambda: test_one_day_every_case3_branch_and_the_last_event()),
    (2, lambda: test_one_day_every_case3_branch_and_the_last_event()),
    (3, lambda: test_event_choice_by_cumulative_chance_and_closed_locations(())),
    (4, lambda: test_event_choice_by_cumulative_chance_and_closed_locations(("S7",))),
    (4, lambda: test_single_event_closed_goes_home_but_work_stays()),
    (5, lambda: test_diagnosis_and_location_quarantine(3)),
    (6, lambda: test_diagnosis_and_location_quarantine(3)),
    (7, lambda: test_planning_school_work_quarantine_and_child_care(4)),
    (8, lambda: test_one_quarantined_classroom_leaves_the_school_open()),
])
def test_pins_fail_on_a_planted_mistake(code, check):
    """Each planted mistake of the oracle (oracle.cpp, `plant`) breaks one rule
    of the movement or planning routine; the pin that covers the rule fails."""
    from oracle.oracle import plant
    plant(code)
    try:
        with pytest.raises(AssertionError):
            check()
    finally:
        plant(0)


def test_planning_lookup_draws_a_workplace_and_takes_the_first_classroom():
    """The planning helpers look locations up with home, school and classroom
    as the first-match types (:1208-1213): an agent with two workplaces has
    one drawn at random, an agent with two classrooms gets the first.
    A (I1) is diagnosed at hospital at 8:00 and, under policy 3, quarantines
    its workplace W1 and its classroom C0 until 1488.  400 workers list W0
    then W1, 400 pupils list C1 then C0 (agent types with weekend schedules
    only: on this Monday they stay at home and meet nobody).  At the next
    midnight about half of the workers draw W1 and are quarantined, and no
    pupil is (its classroom is C1)."""
    shops = [event(SHOP7, 1, 15, 15.1, 0.3), event(SHOP7, 1, 16, 16.1, 0.3)]
    t_a = agent_type(1, [schedule("W", "ALL", [event(HOSPITAL, 1, 8, 8.1, 1)] + shops)])
    t_weekend = agent_type(2, [schedule("W", "WEEKENDS", [event(SHOP7, 1, 10, 10.1, 1)])])
    places = [place("SC0", 3), place("C0_SC0", 33), place("C1_SC0", 33), place("H0", HOME), place("H1", HOME),
              place("W0", WORK), place("W1", WORK), place("S7", SHOP7), place("HOSP", HOSPITAL)]
    n = 400
    people = [person(1, [(HOME, "H0"), (WORK, "W1"), (33, "C0_SC0"), (SHOP7, "S7"), (HOSPITAL, "HOSP")],
                     state="I1")]
    people += [person(2, [(HOME, "H1"), (WORK, "W0"), (WORK, "W1")]) for _ in range(n)]
    people += [person(2, [(HOME, "H1"), (3, "SC0"), (33, "C1_SC0"), (33, "C0_SC0")], age=10) for _ in range(n)]
    w = MiniWorld(places, people, [t_a, t_weekend])
    o = OracleSim(w.args(quarantine_policy=3))
    try:
        o.steps(145)
        q = o.field("quarantined")
        until = o.field("quarantinedUntilTimestamp")
    finally:
        o.close()
    workers, pupils = q[1:n + 1], q[n + 1:]
    assert abs(workers.mean() - 0.5) <= 4.5 * np.sqrt(0.25 / n)
    assert set(until[1:n + 1][workers == 1].tolist()) == {1488}
    assert pupils.sum() == 0
