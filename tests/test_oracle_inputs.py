"""The input layer the oracle shares with the product (csrc/host/json.cpp,
scenario.cpp) pinned against the input files read independently here, in
plain Python:

* the agent-type schedules (agentTypes.json; AgentTypeList,
  agentType.cpp:4-26): every event's location type, chance and its start,
  end and duration converted as TimeDay(double) does (timeDay.cpp:58-62:
  hours = the integer part, minutes = round(fraction / 0.6 * 60), the JSON
  number first cast to float, agentType.cpp:6-11), stored per slot
  28 * type + 7 * WB + day (agentType.cpp:13-15) in slot order;
* the dummy population of -n/-N (DataProvider::randomLocations /
  randomAgents, dataProvider.cpp:102-184) by what its definition fixes:
  N locations plus the cemetery, location-type and agent-type shares equal
  to configRandom's distributions (tourists, type 9, skipped at load,
  agentsList.h:175), ages uniform on 0..89, every agent holding one entry per
  location type its schedules name plus hospital, public space, home and
  doctor, a switched ("irregular") entry only for types with a non-zero
  chanceForType and then, as the reference's loop leaves its found flag
  unset, at a random place of any type, and the initial location at home
  (agentsList.h:251-256).
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

CFG = os.path.join(pansim_amd.DATA_DIR, "inputConfigFiles")
WB_INDEX = {"W": 0, "N": 1, "M": 2, "S": 3}
DAYS = {"ALL": range(7), "WEEKDAYS": range(5), "WEEKENDS": (5, 6), "MONDAY": (0,), "TUESDAY": (1,),
        "WEDNESDAY": (2,), "THURSDAY": (3,), "FRIDAY": (4,), "SATURDAY": (5,), "SUNDAY": (6,)}


def _timeday(x):
    raw = float(np.float32(x))
    if raw == -1.0:
        return 255, 0
    h = int(raw)
    return h, int(np.round((raw - h) / 0.6 * 60)) & 0xFF


def _expected_schedule(path):
    types = json.load(open(path))["types"]
    slots = {}
    for t_idx, t in enumerate(types):
        for su in t["schedulesUnique"]:
            evs = []
            for e in su["schedule"]:
                evs.append((e["type"], np.float32(e["chance"]), *_timeday(e["start"]), *_timeday(e["end"]),
                            *_timeday(e["duration"])))
            for d in DAYS[su["dayType"]]:
                # a later schedule for the same slot is inserted in front of the earlier one (addSchedule)
                slot = 28 * t_idx + 7 * WB_INDEX[su["WB"]] + d
                slots[slot] = evs + slots.get(slot, [])
    offsets = [0]
    events = []
    for slot in range(28 * len(types)):
        events += slots.get(slot, [])
        offsets.append(len(events))
    return offsets, events


@pytest.mark.parametrize("types_file", ["agentTypes.json", "agentTypes_3.json"])
def test_schedule_tables_match_the_agent_types_file(types_file):
    args = pansim_amd.config1_args(200, 50, 1, 0, extra=["-A", f"inputConfigFiles/{types_file}"])
    o = OracleSim(args)
    try:
        off, ty, ch, tm = o.schedule()
    finally:
        o.close()
    exp_off, exp_ev = _expected_schedule(os.path.join(CFG, types_file))
    assert off.tolist() == exp_off
    assert len(ty) == len(exp_ev)
    for k, e in enumerate(exp_ev):
        got = (int(ty[k]), np.float32(ch[k]), *[int(x) for x in tm[k]])
        assert got == e, (k, got, e)


def test_timeday_conversion_examples():
    # the reference's h.mm notation: .3 is 30 minutes, .1 ten, .5 fifty
    assert _timeday(8.3) == (8, 30) and _timeday(10.1) == (10, 10) and _timeday(12.5) == (12, 50)
    assert _timeday(24) == (24, 0) and _timeday(1.3) == (1, 30) and _timeday(-1) == (255, 0)


def _dist(cfg, key):
    return {d["value"]: d["chance"] for d in cfg[key]}


@pytest.fixture(scope="module")
def dummy():
    n, nl = 40_000, 8_000
    o = OracleSim(pansim_amd.config1_args(n, nl, 1, 3))
    cfg = json.load(open(os.path.join(pansim_amd.DATA_DIR, "configs", "configRandom_E0.001.json")))
    yield o, n, nl, cfg
    o.close()


def test_dummy_locations(dummy):
    o, n, nl, cfg = dummy
    assert o.n_locations == nl + 1 and o.cemetery == nl
    lt = o.location_types()[:nl]
    dist = _dist(cfg, "locationTypeDistribution")
    for v, p in dist.items():
        share = float((lt == int(v)).mean())
        # a classroom drawn before the first school becomes a school (dataProvider.cpp:111-120):
        # a few locations at the start at most
        tol = 4.5 * np.sqrt(p * (1 - p) / nl) + 2e-3 * (v in ("3", "33"))
        assert abs(share - p) <= tol, (v, share, p)


def test_dummy_agents(dummy):
    o, n, nl, cfg = dummy
    types = json.load(open(os.path.join(CFG, "agentTypes.json")))["types"]
    dist = _dist(cfg, "agentTypeDistribution")
    p_tourist = dist["9"]
    # tourists skipped: the population is the rest
    assert abs(o.n_agents / n - (1 - p_tourist)) <= 4.5 * np.sqrt(p_tourist * (1 - p_tourist) / n)
    at = o.field("agentType")  # index into agentTypes.json's list
    ids = np.array([t["ID"] for t in types])[at]
    assert not (ids == 9).any()
    for v, p in dist.items():
        if v == "9":
            continue
        q = p / (1 - p_tourist)
        share = float((ids == int(v)).mean())
        assert abs(share - q) <= 4.5 * np.sqrt(q * (1 - q) / o.n_agents), (v, share, q)
    age = o.field("age")
    assert age.min() == 0 and age.max() == 89
    counts = np.bincount(age, minlength=90)
    exp = o.n_agents / 90
    assert np.all(np.abs(counts - exp) <= 5 * np.sqrt(exp))
    sex = o.field("sex")
    assert abs(sex.mean() - 0.5) <= 4.5 * np.sqrt(0.25 / o.n_agents)


def test_dummy_agent_locations(dummy):
    o, n, nl, cfg = dummy
    types = json.load(open(os.path.join(CFG, "agentTypes.json")))["types"]
    required = []
    for t in types:
        s = {12, 1, 2, 14}  # hospital, public space, home, doctor (dataProvider.cpp:67-79)
        for su in t["schedulesUnique"]:
            s |= {e["type"] for e in su["schedule"]}
        required.append(sorted(s))
    off, locs, ptypes = o.csr()
    at = o.field("agentType")
    lt = o.location_types()
    irregular = {d["value"]: d["chanceForType"] for d in cfg["irregularLocationChance"]["detailsOfChances"]}
    switched = {}
    for i in range(0, o.n_agents, 7):
        b, e = int(off[i]), int(off[i + 1])
        assert ptypes[b:e].tolist() == required[at[i]], i
        for k in range(b, e):
            t = int(ptypes[k])
            sw = int(lt[locs[k]]) != t
            switched.setdefault(t, []).append(sw)
    present = set(lt[:nl].tolist())
    for t, sw in switched.items():
        c = irregular.get(str(t), 0.0)
        frac = float(np.mean(sw))
        if c == 0.0 and t in present:
            assert frac == 0.0, t  # only types with a chanceForType are ever switched
        elif c > 0.0 and t in present and len(sw) > 500:
            # a switched entry whose new type has locations takes a random place
            # of any type (the found flag stays unset, dataProvider.cpp:155-176),
            # so it differs from the requested type with chance 1 - share(t)
            p = c * (1.0 - float((lt[:nl] == t).mean()))
            assert abs(frac - p) <= 4.5 * np.sqrt(p * (1 - p) / len(sw)) + 1e-3, (t, frac, p)
    # initial location: the first home entry (agentsList.h:251-256)
    loc0 = OracleSim(pansim_amd.config1_args(n, nl, 1, 3)).field("location")
    for i in range(0, o.n_agents, 97):
        b, e = int(off[i]), int(off[i + 1])
        homes = [int(locs[k]) for k in range(b, e) if ptypes[k] == 2]
        assert loc0[i] == (homes[0] if homes else 0)


def test_synthetic_schedule_jitter():
    """--synthJitter 30 (the default): the synthetic population's agent types
    are the file's nine, then, for every type but the afternoon shift (ID 7,
    whose index the reference's curfew exception tests), six variants whose
    events start and end -30, -20, -10, +10, +20 and +30 minutes from the
    base type's (clamped to the day), with the same location types, chances
    and durations; every agent of a shifted type family follows one of the
    seven schedules with equal chance."""
    o = OracleSim(pansim_amd.with_data_dir(pansim_amd.synthetic_args(30000, 1, 0, 1)))
    try:
        off, ty, ch, tm = o.schedule()
        at = o.field("agentType")
    finally:
        o.close()
    base_types = json.load(open(os.path.join(CFG, "agentTypes.json")))["types"]
    nb = len(base_types)
    shifts = [-30, -20, -10, 10, 20, 30]
    assert len(off) == 28 * (nb + 6 * (nb - 1)) + 1
    mins = lambda h, m: int(h) * 60 + int(m)
    variant = nb
    for b in range(nb):
        if base_types[b]["ID"] == 7:
            continue
        for d in shifts:
            for slot in range(28):
                e0, e1 = off[28 * b + slot], off[28 * b + slot + 1]
                v0, v1 = off[28 * variant + slot], off[28 * variant + slot + 1]
                assert v1 - v0 == e1 - e0
                for k in range(e1 - e0):
                    x, y = e0 + k, v0 + k
                    assert ty[x] == ty[y] and ch[x] == ch[y] and tuple(tm[x][4:]) == tuple(tm[y][4:])
                    for p in (0, 2):
                        want = min(1440, max(0, mins(*tm[x][p:p + 2]) + d))
                        assert mins(*tm[y][p:p + 2]) == want, (b, d, slot, k)
            variant += 1
    # agents of a shifted family spread evenly over the seven schedules
    worker = [t["ID"] for t in base_types].index(6)
    fam = np.concatenate([[worker], nb + 6 * (worker - (1 if worker > 6 else 0)) + np.arange(6)])
    counts = np.array([(at == t).sum() for t in fam])
    assert counts.sum() > 5000
    exp = counts.sum() / 7
    assert np.all(np.abs(counts - exp) <= 4.5 * np.sqrt(exp)), counts
