"""The oracle's closure rules (RuleClosure, closurePolicies.h:84-480) pinned
by what the conditions' definitions fix.  "afterDays"
(closurePolicies.h:117-125): a rule with threshold t and openAfter n is
active on the days d (counted from 0 at the first midnight block) with
t <= d < t + n, and while it is active every location of the listed types
is closed, except the essential ones (closurePolicies.h:425-470), and only
those.  Checked one step after each midnight block over
twelve days, for rules with different windows and type lists, from rule
files written here.  "hospitalizedFraction" (closurePolicies.h:127-165): the
day's fraction of agents in the hospital states (statistics entries whose
state name contains "_h") enters a history of max(closeAfter, openAfter) + 2
days; an inactive rule turns active when the fraction exceeds threshold and
none of the last closeAfter values was below it, an active one turns
inactive when the fraction is at most threshold2 and none of the last
openAfter values reached it; the expected activity is derived from the
run's own statistics vectors and compared with the closed locations; the
same for "newDeadFraction" (closurePolicies.h:166-205), whose value is the
day's new COVID deaths (state D1) as a fraction of the agents."""
import json

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144


@pytest.mark.parametrize("threshold,open_after,types", [(3, 4, [5, 6]), (0, 2, [4]), (5, 1, [7, 44]),
                                                        (2, 6, [3, 33])])
def test_after_days_rule_closes_exactly_its_types_in_its_window(tmp_path, threshold, open_after, types):
    rules = {"rules": [{"name": "window", "conditionType": "afterDays", "threshold": threshold,
                        "threshold2": threshold, "openAfter": open_after, "closeAfter": -1, "parameter": 0,
                        "locationTypesToClose": types}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(5000, 2, 3, 2, extra=["--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        lt = o.location_types()
        listed = np.isin(lt, types) & (o.field("locEssential") != 1)
        assert listed.sum() > 0
        for day in range(12):
            o.steps(day * SPD + 1 - o.timestamp)
            closed = o.field("locOpen") == 0
            active = threshold <= day < threshold + open_after
            if active:
                assert np.array_equal(closed, listed), (day, int(closed.sum()), int(listed.sum()))
            else:
                assert not closed.any(), (day, np.unique(lt[closed]))
    finally:
        o.close()


@pytest.mark.parametrize("kind,thr,thr2,close_after,open_after", [
    ("hospitalizedFraction", 0.02, 0.015, 2, 3), ("newDeadFraction", 0.0008, 0.0004, 2, 2)])
def test_statistics_condition_follows_the_statistics(tmp_path, kind, thr, thr2, close_after, open_after):
    types = [5, 6]
    rules = {"rules": [{"name": "cond", "conditionType": kind, "threshold": thr,
                        "threshold2": thr2, "openAfter": open_after, "closeAfter": close_after, "parameter": 0,
                        "locationTypesToClose": types}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(5000, 6, 3, 2, extra=[
        "-k", "0.05", "--synthInfected", "0.05", "--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        lt = o.location_types()
        listed = np.isin(lt, types) & (o.field("locEssential") != 1)
        names = o.state_names()
        hpos = [i for i, n in enumerate(names) if "_h" in n]
        assert len(hpos) == 3
        dead = names.index("D1")
        prev_dead = 0
        hist = [0.0] * (max(close_after, open_after) + 2)
        pos, active, switches = 0, False, 0
        for day in range(40):
            o.steps(day * SPD + 1 - o.timestamp)
            st = o.day_stats(day)
            if kind == "hospitalizedFraction":
                value = sum(st[i] for i in hpos) / float(o.n_agents)
            else:  # new COVID deaths since the previous midnight
                value = (st[dead] - prev_dead) / float(o.n_agents)
                prev_dead = st[dead]
            hist[pos] = value
            pos = (pos + 1) % len(hist)
            last = lambda k: [hist[(pos - 1 - j) % len(hist)] for j in range(k)]  # noqa: E731
            was = active
            if not active and value > thr:
                active = not any(v < thr for v in last(close_after))
            elif active and value <= thr2:
                active = any(v >= thr2 for v in last(open_after))
            switches += was != active
            closed = o.field("locOpen") == 0
            assert np.array_equal(closed, listed if active else np.zeros_like(listed)), (day, value, active)
        assert switches >= 2  # the rule closed and reopened within the run
    finally:
        o.close()


def test_masks_rule_scales_non_home_infectiousness_in_its_window(tmp_path):
    # Masks (closurePolicies.h:214-245): while active, every non-home
    # location's infectiousness is multiplied by the parameter (a float times
    # a double, stored as float); when it ends, divided by it again
    rules = {"rules": [{"name": "Masks", "conditionType": "afterDays", "threshold": 2, "threshold2": 2,
                        "openAfter": 3, "closeAfter": -1, "parameter": 0.8, "locationTypesToClose": []}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(3000, 2, 3, 2, extra=["--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        home = int(json.load(open(pansim_amd.DATA_DIR + "/inputConfigFiles/locationTypes.json"))["home"])
        non_home = o.location_types() != home
        inf = lambda: o.field("locInfectiousness").view(np.float32).copy()  # noqa: E731
        o.steps(1)
        orig = inf()
        assert (orig[non_home] > 0).any()
        masked = orig.copy()
        masked[non_home] = (orig[non_home].astype(np.float64) * 0.8).astype(np.float32)
        after = masked.copy()
        after[non_home] = (masked[non_home].astype(np.float64) / 0.8).astype(np.float32)
        for day in range(8):
            o.steps(day * SPD + 1 - o.timestamp)
            expect = masked if 2 <= day < 5 else (after if day >= 5 else orig)
            assert np.array_equal(inf(), expect), day
    finally:
        o.close()


def test_expose_to_mutation_infects_the_given_share_of_the_susceptible(tmp_path):
    # ExposeToMutation (closurePolicies.h:347-391): on the rule's day each
    # susceptible agent gets the variant given by threshold2 with probability
    # parameter * susceptibility (1 for every agent in S before any
    # vaccination), so the exposed count is Binomial(S, parameter), every one
    # of them in the first infected state with that variant
    frac, variant, day = 0.05, 1, 3
    rules = {"rules": [{"name": "ExposeToMutation", "conditionType": "afterDays", "threshold": day,
                        "threshold2": variant, "openAfter": 1, "closeAfter": -1, "parameter": frac,
                        "locationTypesToClose": []}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(20000, 2, 3, 2, extra=["--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        o.steps(day * SPD + 1)
        ts = day * SPD
        names = o.state_names()
        susceptible = o.day_stats(day)[names.index("S")]  # taken before the rules ran
        new = o.field("infectedTimestamp") == ts
        var = o.field("variant")
        exposed = new & (var == variant)
        n = int(exposed.sum())
        assert abs(n - frac * susceptible) <= 5.0 * np.sqrt(susceptible * frac * (1 - frac)), (n, susceptible)
        assert n > 0 and np.all(o.field("state")[exposed] == names.index("E"))
        assert not np.any((var == variant) & ~new)  # nobody else carries the variant
    finally:
        o.close()


def test_curfew_sends_everyone_home_at_its_start(tmp_path):
    # Curfew (closurePolicies.h:247-272; doMovement's curfew branch,
    # movementPolicies.h:267-326): at the curfew's first step every living
    # agent not in hospital goes home, except the agents of type 7 (night
    # workers, who may stay at an open workplace); on a day without the
    # curfew some agents are still out at that time
    rules = {"rules": [{"name": "Curfew", "conditionType": "afterDays", "threshold": 1, "threshold2": 1,
                        "openAfter": 3, "closeAfter": -1, "parameter": "2000-0500", "locationTypesToClose": []}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(5000, 2, 3, 2, extra=[
        "-k", "0.05", "--synthInfected", "0.05", "--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        off, locs, types = o.csr()
        n = o.n_agents
        home_t = int(json.load(open(pansim_amd.DATA_DIR + "/inputConfigFiles/locationTypes.json"))["home"])
        home = np.full(n, -1, dtype=np.int64)
        owner = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
        is_home = types == home_t
        first = np.unique(owner[is_home], return_index=True)
        home[first[0]] = locs[is_home][first[1]]
        wb, _ = o.state_table()
        step_20 = 20 * 60 // 10
        for day, curfew in ((1, True), (2, True), (3, True), (5, False)):
            o.steps(day * SPD + step_20 + 1 - o.timestamp)
            ts = day * SPD + step_20
            w = wb[o.field("state").astype(np.int64)]
            hosp = (o.field("hospitalizedTimestamp") <= ts) & (o.field("hospitalizedUntilTimestamp") > ts)
            subject = (w != 3) & (w != 4) & ~hosp & (o.field("agentType") != 6) & (home >= 0)
            away = subject & (o.field("location").astype(np.int64) != home)
            if curfew:
                assert not away.any(), (day, int(away.sum()))
            else:
                assert away.sum() > 0
    finally:
        o.close()
