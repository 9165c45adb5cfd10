"""Replica batches: every replica of an Ensemble (one movement and one
infection launch per step for all replicas) matches the oracle run of its own
command line bit for bit, as it does when run alone."""
import pytest

import pansim_amd
from oracle.oracle import OracleSim
from tests.parity_util import compare_days, compare_state

pytestmark = pytest.mark.gpu


def _replicas():
    return [
        pansim_amd.config1_args(2500, 500, 1, 3, extra=["-k", "0.02", "--quarantinePolicy", "3"]),
        pansim_amd.synthetic_args(5000, 1, 7, 2, extra=["-k", "0.01", "--synthInfected", "0.01"]),
        pansim_amd.config1_args(900, 200, 1, 11, extra=["-k", "0.03", "--quarantinePolicy", "1"]),
        pansim_amd.config1_args(1800, 300, 1, 5, extra=[
            "-k", "0.03", "--enableClosures", "1", "--closures", "inputConfigFiles/closureRules.json"]),
    ]


def test_ensemble_matches_each_replica_alone():
    args = _replicas()
    ens = pansim_amd.Ensemble(args)
    oracles = [OracleSim(a) for a in args]
    try:
        done = 0
        for cp in (1, 150, 300, 1008):
            ens.steps(cp - done)
            for o in oracles:
                o.steps(cp - done)
            done = cp
            for r, o in enumerate(oracles):
                compare_state(ens.replica(r), o, f"replica {r} after {cp} steps")
        for r, o in enumerate(oracles):
            compare_days(ens.replica(r), o)
    finally:
        ens.close()
        for o in oracles:
            o.close()


def test_day_graph_replays_match_each_replica_alone():
    """Whole days as replays of the captured day graph (pansim_batch_run_day):
    the first day eagerly step by step up to a midnight, then six days of
    graph replays with the midnight blocks (eager work and synchronisations)
    between them; after every replay each replica equals its own oracle run
    in every field, and every day's statistics agree."""
    args = _replicas()
    ens = pansim_amd.Ensemble(args)
    oracles = [OracleSim(a) for a in args]
    try:
        ens.steps(144)
        for o in oracles:
            o.steps(144)
        for day in range(1, 7):
            ens.days(1)
            ens.sync()
            for o in oracles:
                o.steps(144)
            for r, o in enumerate(oracles):
                compare_state(ens.replica(r), o, f"replica {r} after the replay of day {day}")
        for r, o in enumerate(oracles):
            compare_days(ens.replica(r), o)
    finally:
        ens.close()
        for o in oracles:
            o.close()


@pytest.mark.parametrize("tiles", [(8, 4), (16, 8), (32, 4)])
def test_fixed_tile_sizes_give_the_same_results(tiles):
    """Every tile size the autotuner may pick gives each replica its own
    oracle run: a day of eager steps and two graph days with the tile sizes
    fixed."""
    args = _replicas()[:3]
    ens = pansim_amd.Ensemble(args)
    oracles = [OracleSim(a) for a in args]
    try:
        ens.set_tiles(*tiles)
        assert ens.tiles() == tiles
        ens.steps(144)
        ens.days(2)
        ens.sync()
        for o in oracles:
            o.steps(3 * 144)
        for r, o in enumerate(oracles):
            compare_state(ens.replica(r), o, f"replica {r}, tiles {tiles}")
    finally:
        ens.close()
        for o in oracles:
            o.close()


def test_autotuner_picks_a_candidate():
    ens = pansim_amd.Ensemble(_replicas()[:2])
    try:
        assert ens.tiles() == (0, 0)
        ens.steps(144)  # tuning: the eager day cycles through the six pairs
        ens.days(1)
        m, i = ens.tiles()
        assert m in (8, 16, 32) and i in (4, 8)
    finally:
        ens.close()
