"""GPU parity: the HIP path against the CPU oracle, bit-exact on every agent
and location field, the per-location infectiousness sums and the per-day
statistics vectors.  Small populations (seconds on the oracle), high
infection coefficients so that every state of the progression, testing,
quarantine and hospital paths is reached within the run."""
import json
import os

import numpy as np
import pytest

import pansim_amd
from tests.parity_util import run_parity

pytestmark = pytest.mark.gpu


def _cfg(n=3000, nl=600, weeks=2, seed=0, k="0.02", extra=()):
    return pansim_amd.config1_args(n, nl, weeks, seed, extra=["-k", k, *extra])


def test_parity_config1_small():
    # midnight, mid-day and end-of-day checkpoints over two weeks
    run_parity(_cfg(), [1, 37, 144, 145, 300, 1000, 2016])


@pytest.mark.parametrize("qp", [0, 1, 2, 4])
def test_parity_quarantine_policies(qp):
    run_parity(_cfg(n=2500, nl=500, weeks=1, seed=qp + 1, extra=["--quarantinePolicy", str(qp)]), [100, 500, 1008])


def test_parity_default_closure_rules_and_antigen():
    args = pansim_amd.config1_args(2500, 500, 1, 7, extra=[
        "-k", "0.03", "--enableClosures", "1", "--closures", "inputConfigFiles/closureRules.json",
        "--testingMethod", "antigen", "--testingProbabilities", "0.01,0.5,0.2,0.2,0.05,0.3"])
    run_parity(args, [144, 700, 1008])


def test_parity_run_for_day_flags():
    args = _cfg(n=2500, nl=500, weeks=1, seed=11, k="0.03")
    days = [("day", []), ("day", ["CF2000-0500", "TP015"]), ("day", ["PL0", "SO12", "MA0.8"]),
            ("day", ["QU2", "TP035"]), ("day", ["CFNONE", "PLNONE", "MA1.0", "QU3", "SONONE"]), ("day", [])]
    run_parity(args, days)


def test_parity_synthetic_population():
    args = pansim_amd.synthetic_args(6000, 1, 3, 5, extra=["-k", "0.01", "--synthInfected", "0.01"])
    run_parity(args, [60, 144, 600, 1008])


def test_parity_mutation_and_rules(tmp_path):
    rules = {"rules": [
        {"name": "ExposeToMutation", "conditionType": "afterDays", "threshold": 1, "threshold2": 2,
         "parameter": "0.02", "openAfter": 2, "closeAfter": -1, "locationTypesToClose": []},
        {"name": "Masks", "conditionType": "afterDays", "threshold": 2, "threshold2": 0,
         "parameter": "0.6", "openAfter": 3, "closeAfter": -1, "locationTypesToClose": []},
        {"name": "Curfew", "conditionType": "afterDays", "threshold": 3, "threshold2": 0,
         "parameter": "2100-0600", "openAfter": 2, "closeAfter": -1, "locationTypesToClose": []},
        {"name": "schools", "conditionType": "afterDays", "threshold": 1, "threshold2": 0,
         "parameter": "0", "openAfter": 3, "closeAfter": -1, "locationTypesToClose": [3, 33, 8]},
        {"name": "ReduceMovement", "conditionType": "afterDays", "threshold": 2, "threshold2": 0,
         "parameter": "0.3", "openAfter": 2, "closeAfter": -1, "locationTypesToClose": []},
        {"name": "HolidayMode", "conditionType": "afterDays", "threshold": 5, "threshold2": 0,
         "parameter": "0", "openAfter": 1, "closeAfter": -1, "locationTypesToClose": []},
        {"name": "hosp", "conditionType": "hospitalizedFraction", "threshold": 0.0001, "threshold2": 0.00005,
         "parameter": "0", "openAfter": 2, "closeAfter": 1, "locationTypesToClose": [5, 6]},
    ]}
    p = tmp_path / "rules.json"
    p.write_text(json.dumps(rules))
    args = pansim_amd.config1_args(2500, 500, 1, 21, extra=["-k", "0.03", "--closures", str(p), "--enableClosures", "1"])
    run_parity(args, [144, 300, 600, 1008])


def test_parity_vaccination_and_boosters():
    # first doses from day 2, boosters from day 125 (immunization.h:62-900):
    # every field, including the category rounds and immunization stats
    args = pansim_amd.config1_args(1500, 300, 19, 4, extra=[
        "-k", "0.005", "--immunizationStart", "2", "--immunizationsPerWeek", "700", "--boosterStart", "125"])
    run_parity(args, [289, 1008, 126 * 144 + 1, 19 * 1008])


@pytest.mark.parametrize("policy", ["dummy", "none"])
def test_parity_dummy_and_no_movement(policy):
    # DummyMovement / NoMovement (movementPolicies.h:9-65) through
    # --movementPolicy: locations, occupancy, sums and every statistic
    args = _cfg(n=2500, nl=400, weeks=1, seed=13, k="0.05", extra=["--movementPolicy", policy])
    run_parity(args, [1, 30, 144, 200, 1008], countdown=False)


@pytest.mark.parametrize("n,infected", [(1, 0.3), (7, 0.3), (63, 0.3), (65, 0.3), (2049, 0.3), (8193, 0.3),
                                        (2049, 0.0), (2049, 1.0)])
def test_parity_tiny_ragged_and_extreme_populations(n, infected):
    # populations smaller than one wave or one tile, and one agent past a
    # movement tile (8192) or an infection tile (2048), so that the padded
    # and partly filled tiles are exercised; nobody infected (no infection
    # pass at all) and everybody infected at the start
    args = pansim_amd.synthetic_args(n, 1, 3, 5, extra=["-k", "0.3", "--synthInfected", str(infected)])
    run_parity(args, [1, 70, 144, 145, 300])


@pytest.mark.parametrize("flags", [["--testingPolicy", "none"], ["--closurePolicy", "none"],
                                   ["--testingPolicy", "none", "--closurePolicy", "none"]])
def test_parity_no_testing_and_no_closure(flags):
    # NoTesting / NoClosure (testingPolicies.h:11-20, closurePolicies.h:36-46)
    # through --testingPolicy / --closurePolicy, with the shipped closure rules
    args = pansim_amd.synthetic_args(3000, 1, 9, 5, extra=["-k", "0.02", "--synthInfected", "0.02", *flags])
    run_parity(args, [1, 144, 145, 600, 1008])


@pytest.mark.parametrize("dt", [5, 15, 30, 60])
def test_parity_other_time_steps(dt):
    # the step length (-t) sets the steps per day, the schedule window of
    # each step, the step count of every stay and quarantine, and the
    # midnight cadence; checkpoints after the first step, mid-morning, at
    # and after midnight, mid-day 5 and at the end of two weeks
    spd = 1440 // dt
    args = _cfg(n=2500, nl=500, weeks=2, seed=20 + dt, k="0.03", extra=["-t", str(dt)])
    run_parity(args, [1, spd // 3, spd, spd + 1, 5 * spd + spd // 2, 14 * spd])


def test_parity_synthetic_population_15_minute_steps():
    # the generator's schedule jitter moves events by multiples of the step
    args = pansim_amd.synthetic_args(6000, 1, 4, 6, extra=["-k", "0.01", "--synthInfected", "0.01", "-t", "15"])
    run_parity(args, [40, 96, 97, 4 * 96 + 50, 7 * 96])


def _long_list_world(tmp_path, n=1500, seed=3):
    """Agents whose possible-location lists hold 20 to 40 entries (two to four
    of every type the shipped schedules visit), half of them, and short
    lists for the rest: lists longer than 16 entries take the engine's
    in-memory lookup, the short ones its register copy, in the same waves."""
    from tests.mini_world import person, place
    rng = np.random.default_rng(seed)
    places, kinds = [], {}

    def add(t, count, essential=0):
        ids = []
        for _ in range(count):
            ids.append(len(places))
            places.append(place(f"L{len(places)}", t, essential=essential))
        kinds[t] = ids

    add(2, n // 2)                      # homes
    add(4, n // 12)                     # workplaces
    add(3, 6)                           # schools
    kinds[33] = []                      # classrooms, five per school ("class_school" ids)
    for j in range(30):
        kinds[33].append(len(places))
        places.append(place(f"C{j}_L{kinds[3][j % 6]}", 33))
    for t in (1, 5, 6, 7, 8, 9, 10, 15, 44):
        add(t, 25)
    add(14, 6, essential=1)             # doctors
    add(12, 2, essential=1)             # hospitals
    people = []
    for i in range(n):
        cls = int(rng.integers(0, 30))
        locs = [(2, kinds[2][i // 2]), (4, int(rng.choice(kinds[4]))), (3, kinds[3][cls % 6]), (33, kinds[33][cls]),
                (12, int(rng.choice(kinds[12]))), (14, int(rng.choice(kinds[14])))]
        per = 2 + int(rng.integers(0, 3)) if i % 2 == 0 else 1
        for t in (1, 5, 6, 7, 8, 9, 10, 15, 44):
            for _ in range(per if i % 2 == 0 else int(rng.integers(0, 2))):
                locs.append((t, int(rng.choice(kinds[t]))))
        if i % 2 == 0:
            rng.shuffle(locs)
        state = "E" if rng.random() < 0.03 else "S"
        people.append(person(int(rng.integers(1, 9)), [(t, places[l]["ID"]) for t, l in locs], state=state,
                             age=int(rng.integers(3, 90)), sex="M" if i % 3 else "F"))
    (tmp_path / "locations.json").write_text(json.dumps({"places": places}))
    (tmp_path / "agents.json").write_text(json.dumps({"people": people}))
    return ["-a", str(tmp_path / "agents.json"), "-l", str(tmp_path / "locations.json"), "-w", "2", "-t", "10",
            "-k", "0.05", "--quarantinePolicy", "3", "--closures", "inputConfigFiles/closureRules.json",
            "--seed", str(seed)]


def test_parity_long_location_lists(tmp_path):
    run_parity(_long_list_world(tmp_path), [1, 50, 144, 145, 700, 1500, 2016])


def test_parity_reference_example_population(tmp_path):
    # the reference's own example input (tests/real_example.py): real
    # schedules' location types, lists of about 19 entries, 1 % exposed at
    # the start, two weeks with the shipped closure rules
    from tests import real_example
    run_parity(real_example.args(tmp_path), [1, 48, 144, 145, 700, 1500, 2016])


def test_parity_config1_full_size():
    # the CPU-runnable workload at its size: -n 100000 -N 20000 (96,786 agents
    # after the tourists are skipped, 20,001 locations), two weeks, seed 0;
    # every field and every day's statistics at mid-day, midnight and
    # activity-change checkpoints and at the end
    run_parity(pansim_amd.config1_args(), [1, 48, 145, 500, 1008, 1009, 1500, 2016])
