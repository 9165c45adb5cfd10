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
