"""The OpenMP build of the oracle (liboracle_omp.so: the reference's OpenMP
structure, movementPolicies.h:1717-1719, util.cpp:55-110, util.h:47-103, used
as bench.py's CPU baseline) gives the sequential oracle's results bit for
bit: the random numbers are addressed by (agent, step), each thread keeps its
location writes in agent order, and every location's sum is taken over its
agents in ascending order."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys
sys.path.insert(0, {root!r})
import numpy as np
import pansim_amd
from oracle.oracle import OracleSim
args = pansim_amd.with_data_dir(pansim_amd.synthetic_args(6000, 1, 3, 5, extra=["-k", "0.02", "--synthInfected", "0.02"]))
a, b = OracleSim(args), OracleSim(args, openmp=True)
for cp in (37, 144, 145, 700, 1008):
    a.steps(cp - a.timestamp); b.steps(cp - b.timestamp)
    for f in ("state", "location", "stepsUntilMove", "quarantined", "diagnosed", "infectedTimestamp", "daysInQuarantine",
              "locCount", "locQuarantineUntil", "quarantinedUntilTimestamp", "days"):
        assert np.array_equal(a.field(f), b.field(f)), (cp, f)
    for v in range(len(a.variant_multipliers())):
        assert np.array_equal(a.location_sums(v), b.location_sums(v)), (cp, v)
assert [a.day_stats(d) for d in range(a.num_days)] == [b.day_stats(d) for d in range(b.num_days)]
print("ok")
"""


@pytest.mark.parametrize("threads", ["1", "4"])
def test_openmp_build_equals_the_sequential_oracle(threads):
    env = dict(os.environ, OMP_NUM_THREADS=threads)
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT)], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
