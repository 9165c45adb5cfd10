"""The oracle's output files (CPU): the --outAgentStat JSON
(agentStatOutput.cpp:5-44, rapidjson PrettyWriter with 4-space indentation)
and the --dumpLocationInfectiousList dump (infectionPolicies.h:27-85,
262-441), checked against the oracle's own per-agent fields and against a
layout the mathematics of the format fixes (Python's json module with
indent=4 writes the same text as a 4-space PrettyWriter for these
documents), so that a swapped key, a wrong field or a wrong order fails."""
from __future__ import annotations

import json

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

UMAX = 0xFFFFFFFF


def _run(tmp_path, days=3):
    js, dump = tmp_path / "agents.json", tmp_path / "infectious.txt"
    args = pansim_amd.synthetic_args(3000, 1, 2, 4, extra=[
        "-k", "0.02", "--synthInfected", "0.03", "--outAgentStat", str(js),
        "--dumpLocationInfectiousList", str(dump), "--diags", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    o.steps(days * 144)
    diags = o.finalize()
    return o, js, dump, diags


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    return _run(tmp_path_factory.mktemp("outputs"))


def test_agent_stat_json_layout_and_values(run):
    o, js, _, _ = run
    text = js.read_text()
    doc = json.loads(text)
    # a 4-space PrettyWriter and json.dumps(indent=4) agree on this document
    assert text == json.dumps(doc, indent=4)
    st = doc["Statistics"]
    n = o.n_agents
    assert [a["ID"] for a in st] == list(range(n))
    keys = ["ID", "infectionTime", "infectedCount", "variant", "InfectionLoc", "diagnosisTime", "quarantinedTime",
            "quarantinedUntilTime", "daysInQuarantine", "hospitalizedTime", "hospitalizedUntilTime",
            "immunizationTimestamp", "immunizationCount", "recoveredTimestamp", "worstState"]
    assert list(st[0].keys()) == keys
    fields = {"infectionTime": "infectedTimestamp", "infectedCount": "infectedCount", "variant": "statsVariant",
              "InfectionLoc": "infectedLocation", "diagnosisTime": "diagnosedTimestamp",
              "quarantinedTime": "quarantinedTimestamp", "quarantinedUntilTime": "quarantinedUntilTimestamp",
              "daysInQuarantine": "daysInQuarantine", "hospitalizedTime": "hospitalizedTimestamp",
              "hospitalizedUntilTime": "hospitalizedUntilTimestamp",
              "immunizationTimestamp": "immunizationTimestamp", "immunizationCount": "immunizationCount",
              "recoveredTimestamp": "recoveredTimestamp"}
    for key, field in fields.items():
        np.testing.assert_array_equal(np.array([a[key] for a in st], dtype=np.uint64),
                                      o.field(field).astype(np.uint64), err_msg=key)
    names = o.state_names()
    worst = o.field("worstState").astype(np.int64)
    begin, end = o.field("worstStateTimestamp"), o.field("worstStateEndTimestamp")
    for i in range(n):
        w = st[i]["worstState"]
        assert w["name"] == names[worst[i]]
        assert w["begin"] == begin[i]
        assert w["end"] == (-1 if end[i] == 0 else end[i])
    # the run infected someone, so the checks above saw non-default values
    assert (o.field("infectedTimestamp") != UMAX).sum() > 0


def _dump_arrays(path):
    lines = path.read_text().split("\n")
    count = int(lines[0])
    ints = [np.array(ln.split(), dtype=np.int64) for ln in lines[1:9]]
    vals = np.array(lines[9].split(), dtype=np.float64)
    return count, ints, vals


def test_infectious_list_dump_consistency(run):
    o, _, dump, _ = run
    count, (ts, var, locOff, locIds, newOff, new, infOff, inf), infness = _dump_arrays(dump)
    assert count == len(ts) == len(var) > 0
    assert len(locOff) == len(newOff) == count + 1 and locOff[0] == newOff[0] == 0
    assert len(infOff) == len(locIds) + 1 and infOff[0] == 0
    assert np.all(np.diff(locOff) > 0) and np.all(np.diff(newOff) > 0) and np.all(np.diff(infOff) >= 0)
    assert locOff[-1] == len(locIds) and newOff[-1] == len(new) and infOff[-1] == len(inf) == len(infness)
    infTs, infLoc = o.field("infectedTimestamp"), o.field("infectedLocation")
    loc_of = {}
    for q in range(count):
        ids = locIds[locOff[q]:locOff[q + 1]]
        assert np.all(np.diff(ids) > 0)  # ascending, unique
        agents = new[newOff[q]:newOff[q + 1]]
        # every agent infected in this pass (one short run: no reinfection)
        assert np.all(infTs[agents] == ts[q])
        locs = infLoc[agents]
        assert set(locs.tolist()) == set(ids.tolist())
        order = np.lexsort((agents, locs))
        assert np.array_equal(order, np.arange(len(agents)))  # by location, then agent
        for a in agents:
            loc_of[int(a)] = int(infLoc[a])
    # every infection of the run is in exactly one insert
    infected = np.nonzero((infTs != UMAX) & (infTs >= ts.min()))[0]
    assert sorted(loc_of) == sorted(infected.tolist()) and len(new) == len(infected)
    # the infectious lists: ascending agents per location, positive values of
    # the form infectiousness(state) x multiplier(variant), the multipliers
    # read off the agents infectious at the end
    table = o.state_infectiousness_values().astype(np.float64)
    state, now = o.field("state").astype(np.int64), o.agent_infectious().astype(np.float64)
    live = table[state] > 0
    muls = set(np.round(now[live] / table[state][live], 6).tolist())
    allowed = [t * m for t in table[table > 0] for m in muls]
    assert np.all(infness > 0)
    assert all(any(abs(v - a) <= 1e-5 * a for a in allowed) for v in set(infness.tolist()))
    for j in range(len(locIds)):
        seg = inf[infOff[j]:infOff[j + 1]]
        assert np.all(np.diff(seg) > 0)


def test_diags_summary(run):
    o, _, _, diags = run
    lines = diags.split("\n")
    assert lines[0] == "Total number of days in quarantine for agents 20<age<=65"
    days = int(lines[1])
    assert 0 <= days <= int(o.field("daysInQuarantine").astype(np.uint64).sum())
    ids = [int(x) for x in lines[2].split()]
    counts = [int(x) for x in lines[3].split()]
    assert len(ids) == len(counts)
    used = [c for c in counts if c]
    assert sum(counts) == o.n_agents
    k = len(used)
    assert counts[k:] == [0] * (len(counts) - k) and ids[k:] == [0] * (len(ids) - k)
    assert ids[:k] == sorted(set(ids[:k]))
    # the never-infected agents count at the cemetery's type
    lt = o.location_types()
    never = int((o.field("infectedTimestamp") == UMAX).sum())
    assert counts[ids[:k].index(int(lt[o.cemetery]))] >= never
