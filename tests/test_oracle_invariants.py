"""Invariants of the oracle's simulation that hold whatever the random draws
(CPU only): they catch a wrong index, a dropped term or a misplaced write in
the oracle before it is used as the GPU path's checker.

- every agent is at one of its own possible locations, or at the cemetery
  (movementPolicies.h:69-122, :320-330), and only dead agents are there;
- per-location occupancy and infectiousness sums equal a brute-force recount
  of the agents (util.h:47-103 defines them as the per-location reduction);
- the first statistics entries are the per-state agent counts
  (simulation.h:473-490), the cumulative-infected entry counts agents with a
  recorded infection;
- identical seeds give identical runs; different seeds give different runs.
"""
from __future__ import annotations

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

WB_D = 4


def _args(seed=0, n=3000, nl=600, k="0.02", extra=()):
    return pansim_amd.config1_args(n, nl, 1, seed, extra=["-k", k, *extra])


def _check_locations(o: OracleSim):
    off, locs, _ = o.csr()
    loc = o.field("location")
    cem = o.cemetery
    n = o.n_agents
    owner = np.repeat(np.arange(n, dtype=np.int64), np.diff(off.astype(np.int64)))
    pairs = set(zip(owner.tolist(), locs.tolist()))
    bad = [i for i in range(n) if loc[i] != cem and (i, int(loc[i])) not in pairs]
    assert not bad, f"agents away from their possible locations: {bad[:10]}"
    wb, _ = o.state_table()
    state = o.field("state")
    at_cem = loc == cem
    assert np.all(wb[state[at_cem]] == WB_D), "living agent at the cemetery"


def _check_sums(o: OracleSim):
    loc = o.field("location").astype(np.int64)
    cnt = np.bincount(loc, minlength=o.n_locations)
    assert np.array_equal(cnt, o.field("locCount").astype(np.int64))
    x = o.agent_infectious().astype(np.float64)
    var = o.field("variant").astype(np.int64)
    for v in range(int(var.max()) + 1):
        q = np.floor(x * 4294967296.0).astype(np.int64) * (var == v)
        ref = np.bincount(loc, weights=None, minlength=o.n_locations) * 0
        np.add.at(ref, loc, q)
        assert np.array_equal(ref, o.location_sums(v)), f"variant {v} sums"


def _check_stats(o: OracleSim):
    """Called one step after a midnight block: the statistics were taken at
    its start, and only that step's infections (susceptible -> first
    infected state) changed any agent's state since."""
    ts = o.timestamp
    assert ts % 144 == 1
    st = o.day_stats(o.num_days - 1)
    wb, _ = o.state_table()
    ns = len(wb)
    state = o.field("state").astype(np.int64)
    counts = np.bincount(state, minlength=ns)
    new = o.field("infectedTimestamp") == ts - 1
    first = np.unique(state[new])
    assert len(first) <= 1
    before = counts.copy()
    if new.any():
        np.subtract.at(before, state[new], 1)
        # the newly infected came from the susceptible state(s); the single-variant
        # run here has one: state 0 (S)
        before[0] += int(new.sum())
    assert st[:ns] == before.tolist()
    assert sum(st[:ns]) == o.n_agents
    ever = int((o.field("infectedCount") > 0).sum()) - int(new.sum())
    # layout after the states (simulation.h:491-665): hospitalized, 3 test counters,
    # 3 quarantine counters, max(1, V-1) variant shares, stayed home, immunized today,
    # new infections, cumulative infected, reinfections, boosters, V immunity sums,
    # health-care infected %, health-care new, V variant counts, infected in hospital,
    # vaccinated never infected: 16 + 3V entries for V > 1, 20 for V = 1
    nv = (len(st) - ns - 16) // 3
    assert len(st) - ns == (20 if nv == 1 else 16 + 3 * nv)
    base = ns + 7 + max(1, nv - 1)
    assert st[base + 3] == ever
    it = o.field("infectedTimestamp")
    assert st[base + 2] == int(((it > ts - 1 - 144) & (it <= ts - 1)).sum()) - int(new.sum())


def test_oracle_invariants_over_a_week():
    o = OracleSim(_args(seed=5))
    try:
        _check_locations(o)
        _check_sums(o)
        for target in (1, 40, 144, 145, 290, 600, 865):
            o.steps(target - o.timestamp)
            _check_locations(o)
            _check_sums(o)
        _check_stats(o)
        # the epidemic ran: infections and deaths/hospital paths were reached
        assert (o.field("infectedCount") > 0).sum() > 30
    finally:
        o.close()


def test_oracle_invariants_synthetic_population():
    args = pansim_amd.synthetic_args(4000, 1, 2, 3, extra=["-k", "0.01", "--synthInfected", "0.01"])
    o = OracleSim(args)
    try:
        for target in (1, 144, 500, 577):
            o.steps(target - o.timestamp)
            _check_locations(o)
            _check_sums(o)
        _check_stats(o)
    finally:
        o.close()


def test_oracle_determinism():
    a, b, c = OracleSim(_args(seed=9)), OracleSim(_args(seed=9)), OracleSim(_args(seed=10))
    try:
        for o in (a, b, c):
            o.steps(400)
        for f in ("state", "location", "infectedTimestamp", "stepsUntilMove", "diagnosed"):
            assert np.array_equal(a.field(f), b.field(f)), f
        assert any(not np.array_equal(a.field(f), c.field(f)) for f in ("state", "location", "infectedTimestamp"))
        assert a.day_stats(a.num_days - 1) == b.day_stats(b.num_days - 1)
    finally:
        for o in (a, b, c):
            o.close()


def test_oracle_run_for_day_returns_day_statistics():
    o = OracleSim(_args(seed=3))
    try:
        st = o.run_for_day([])
        assert o.timestamp == 144
        assert st == o.day_stats(o.num_days - 1)
    finally:
        o.close()


@pytest.mark.parametrize("flag", ["XX12", "TPabc", "PL", "QU9"])
def test_oracle_ignores_unknown_flags(flag):
    # processFlags (simulation.h:779-897) acts only on the listed flag values
    a, b = OracleSim(_args(seed=3)), OracleSim(_args(seed=3))
    try:
        assert a.run_for_day([flag]) == b.run_for_day([])
        assert np.array_equal(a.field("state"), b.field("state"))
    finally:
        a.close()
        b.close()


def test_oracle_vaccination_rollout():
    # first doses from day 2 (100 per day), boosters from day 125 (default
    # weekly booster table); immunization.h:62-89, 696-900
    args = pansim_amd.config1_args(1500, 300, 19, 4, extra=[
        "-k", "0.005", "--immunizationStart", "2", "--immunizationsPerWeek", "700", "--boosterStart", "125"])
    o = OracleSim(args)
    try:
        o.steps(2 * 144)
        assert (o.field("immunizationCount") == 0).all()
        o.steps(1)  # midnight of day 2: the first doses
        it = o.field("immunizationTimestamp")
        ic = o.field("immunizationCount")
        rnd = o.field("vaccinationRound")
        first = ic == 1
        assert 0 < first.sum() <= 200
        assert (it[first] == 288).all()
        # only willing agents of a category with a round get a dose
        assert ((rnd[first] >= 1) & (rnd[first] <= 10)).all()
        wb, _ = o.state_table()
        ns = len(wb)
        st = o.day_stats(o.num_days - 1)
        nv = (len(st) - ns - 16) // 3
        vac = st[ns + 7 + max(1, nv - 1) + 1]
        assert vac == int((it == 288).sum())
        # boosters: only on or after day 125, only for agents with a first dose
        o.steps(19 * 1008 - o.timestamp)
        it = o.field("immunizationTimestamp")
        ic = o.field("immunizationCount")
        boosted = ic >= 2
        assert boosted.sum() > 0
        assert (it[boosted] >= 125 * 144).all()
    finally:
        o.close()


def test_invariants_on_the_reference_example_population(tmp_path):
    # the reference's example input (tests/real_example.py): the locations,
    # occupancy and sums invariants at an activity change and at the end of
    # a week
    from tests import real_example
    o = OracleSim(real_example.args(tmp_path, weeks=1, k="0.002"))
    for step in (52, 1008 - 52):
        o.steps(step)
        _check_locations(o)
        _check_sums(o)
