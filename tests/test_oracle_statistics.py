"""The oracle's per-day statistics vector (refreshAndPrintStatistics,
simulation.h:473-665; getQuarantineStats, agentsList.h:55-78; the testing
counts of performTests, testingPolicies.h:459-474; immunizedToday,
immunization.h:897-898) recounted entry by entry from the agents' fields as
they were when the vector was computed (the oracle's statistics snapshot),
on a run where every entry moves: infections with several variants (the
shipped closure rules expose agents to a new variant), testing, quarantine,
vaccination and boosters from early days, non-COVID hospital stays.

Layout: the per-state counts, H (alive, not in a hospital state, in a
non-COVID stay), T, P1, P2 (tested now; diagnosed now; diagnosed during the
last day before now), Q, QT, NQ (quarantined now; of those infected; not
quarantined but infected), MUT (per variant 1.. the percentage of the
infected carrying it, truncated), HOM (stayed home all day minus the two
dead states, entries 10 and 11), VAC, NI (infected during the last day),
INF (ever infected), REINF, BSTR, IMM (per variant the truncated sum of
1 - susceptibility), HCI (percentage of healthcare workers infected or
quarantined), HCE (healthcare workers infected during the last day), INFV
(per variant, agents ever infected with it), INFH (infected agents at a
hospital), VNI (vaccinated, never infected)."""
from __future__ import annotations

import json

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144


def _f32(u):
    return np.asarray(u, dtype=np.uint32).view(np.float32)


def _recount(o, ts, wb, infected_states, hospital_type, lt, n_var):
    g = lambda k: o.field("snap:" + k).astype(np.int64)
    state = g("state")
    ns = len(wb)
    infected = np.isin(state, infected_states)
    out = np.bincount(state, minlength=ns).tolist()
    w = wb[state]
    out.append(int(((w != 4) & (w != 3) & (ts < g("hospitalizedUntilTimestamp"))).sum()))
    diag = g("diagnosedTimestamp")
    if ts > 0:
        out += [int((g("lastTest") == ts).sum()), int((diag == ts).sum()),
                int(((diag < ts) & (diag > (ts - SPD) % 2 ** 32)).sum())]
    else:
        out += [0, 0, 0]
    q = (g("quarantinedTimestamp") <= ts) & (g("quarantinedUntilTimestamp") > ts)
    out += [int(q.sum()), int((q & infected).sum()), int((~q & infected).sum())]
    var = g("variant")
    if n_var > 1:
        allinf = int(infected.sum())
        for v in range(1, n_var):
            c = int((infected & (var == v)).sum())
            out.append(int(c / allinf * 100.0) if allinf else 0)
    else:
        out.append(0)
    out.append(int(g("stayedHome").sum()) - out[10] - out[11])
    # immunization.h:484-485: the pass counts timestamp 0 as 1
    out.append(int((g("immunizationTimestamp") == max(ts, 1)).sum()))
    its = g("infectedTimestamp")
    recent = (its > (ts - SPD) % 2 ** 32) & (its <= ts)  # unsigned difference, as in the reference
    cnt, imm = g("infectedCount"), g("immunizationCount")
    out += [int(recent.sum()), int((cnt > 0).sum()), int(np.maximum(cnt - 1, 0).sum()),
            int(np.maximum(imm - 1, 0).sum())]
    for v in range(n_var):
        s = _f32(o.field(f"snap:susceptible{v}")).astype(np.float64)
        out.append(int(np.floor((1.0 - s.astype(np.float32)).astype(np.float64).sum())))
    hcw = g("healthcareWorker").astype(bool)
    out.append(int((hcw & (infected | (g("quarantined") == 1))).sum()) * 100 // max(1, int(hcw.sum())))
    out.append(int((hcw & recent).sum()))
    sv = g("statsVariant")
    for v in range(n_var):
        out.append(int(((sv >> v) & 1).sum()))
    loc = g("location")
    out.append(int((infected & (lt[loc] == hospital_type)).sum()))
    out.append(int(((cnt == 0) & (imm > 0)).sum()))
    return out


def test_statistics_vector_equals_a_recount_of_the_fields(tmp_path):
    # a new variant on day 3 (ExposeToMutation), vaccination from day 2; boosters
    # need a first dose four months old (immunization.h), so the run is 19 weeks
    rules = {"rules": [{"name": "ExposeToMutation", "conditionType": "afterDays", "threshold": 3, "threshold2": 3,
                        "openAfter": 1, "closeAfter": -1, "parameter": "0.002", "locationTypesToClose": []}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(6000, 19, 4, 2, extra=[
        "-k", "0.004", "--synthInfected", "0.02", "--immunizationStart", "2", "--boosterStart", "100",
        "--closures", str(path)])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        o.snapshot_stats(True)
        wb, _ = o.state_table()
        names = o.state_names()
        infected_states = [names.index(x) for x in ("E", "I1", "I2", "I3", "I4", "I5_h", "I6_h")]
        lt = o.location_types().astype(np.int64)
        moved = {k: 0 for k in ("H", "T", "P1", "P2", "Q", "MUT", "VAC", "BSTR", "IMM", "HCI", "INFH", "VNI")}
        for day in list(range(21)) + list(range(24, 133, 6)):
            o.steps(day * SPD + 1 - o.timestamp)
            ts = day * SPD
            assert int(o.field("snap:timestamp")[0]) == ts
            st = o.day_stats(day)
            n_var = len(o.variant_multipliers())
            exp = _recount(o, ts, wb, infected_states, 12, lt, n_var)
            assert st == exp, (day, [(i, a, b) for i, (a, b) in enumerate(zip(st, exp)) if a != b])
            h = 12
            for k, i in (("H", h), ("T", h + 1), ("P1", h + 2), ("P2", h + 3), ("Q", h + 4)):
                moved[k] += st[i] > 0
            moved["MUT"] += any(st[h + 7:h + 7 + n_var - 1])
            base = h + 7 + n_var - 1
            moved["VAC"] += st[base + 1] > 0
            moved["BSTR"] += st[base + 5] > 0
            moved["IMM"] += st[base + 6] > 0
            moved["HCI"] += st[base + 6 + n_var] > 0
            moved["INFH"] += st[base + 8 + 2 * n_var] > 0
            moved["VNI"] += st[base + 9 + 2 * n_var] > 0
    finally:
        o.close()
    # every entry family took non-zero values on some day
    assert all(v > 0 for v in moved.values()), moved


@pytest.mark.parametrize("s", [0.0, 1.0, 0.3, 0.7, 0.5, 2.0 ** -30, 1 - 2.0 ** -24, 0.25 + 2.0 ** -26])
def test_immunity_terms_are_multiples_of_2_pow_minus_24(s):
    """The IMM sum is kept exactly in 2^-24 fixed point: for a susceptibility
    s in [0, 1] the float 1 - s is a multiple of 2^-24 (s >= 1/2: s itself
    is, and the subtraction is exact; s < 1/2: the result lies in (1/2, 1],
    where every float is)."""
    t = np.float32(1.0) - np.float32(s)
    x = float(t) * 2.0 ** 24
    assert x == int(x)


def test_immunity_terms_dense():
    s = np.random.default_rng(0).random(2_000_000).astype(np.float32)
    t = (np.float32(1.0) - s).astype(np.float64) * 2.0 ** 24
    assert np.array_equal(t, np.floor(t))
