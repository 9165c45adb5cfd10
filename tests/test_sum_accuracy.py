"""The per-location infectiousness sums against the reference's own
summation (util.h:77-82: per location, float additions from 0.0f in
ascending agent id) and against the exact (FP64) sum, on the synthetic
population whose location sizes are those of the 10M-agent workload
(households of 1-6, classrooms of 25, schools of 300-800, lognormal
workplaces capped at 2000).  The engine keeps each sum exactly in 2^-32
fixed point and converts it once to float (the device sums equal the
oracle's bit for bit, GPU parity tests), so its relative difference from the
FP64 sum is at most 2^-24; from the reference's FP32 order the allowed
relative difference is 1e-5.  The measured maxima are printed."""
import numpy as np

import pansim_amd
from oracle.oracle import OracleSim


def test_location_sums_against_fp32_order_and_fp64():
    args = pansim_amd.synthetic_args(100_000, 2, 1, 4, extra=["-k", "0.004", "--synthInfected", "0.05"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        o.steps(7 * 144 + 60)  # 10:00 on day 7: the epidemic is under way
        loc = o.field("location").astype(np.int64)
        var = o.field("variant").astype(np.int64)
        x = o.agent_infectious()
        fixed = o.location_sums(0)
        count = o.field("locCount").astype(np.int64)
    finally:
        o.close()
    on = (var == 0) & (x != 0)
    assert on.sum() > 1000
    exact = fixed.astype(np.float64) * 2.0 ** -32
    device = exact.astype(np.float32)  # the engine's one conversion of the fixed-point sum
    fp32 = np.zeros(len(fixed), dtype=np.float32)
    np.add.at(fp32, loc[on], x[on])  # unbuffered: each location in ascending agent order, float adds
    fp64 = np.zeros(len(fixed))
    np.add.at(fp64, loc[on], x[on].astype(np.float64))
    assert np.array_equal(exact, fp64)  # the fixed-point sum is the exact sum
    nz = fp64 > 0
    rel_fp64 = np.abs(device[nz].astype(np.float64) - fp64[nz]) / fp64[nz]
    rel_fp32 = np.abs(device[nz].astype(np.float64) - fp32[nz].astype(np.float64)) / fp64[nz]
    big = nz & (count >= 200)
    print(f"locations with infectious visitors: {nz.sum()}, largest occupancy there {count[nz].max()}; "
          f"max relative difference vs FP64 {rel_fp64.max():.3e}, vs the ascending-id FP32 sum "
          f"{rel_fp32.max():.3e} (at occupancy >= 200: {(np.abs(device[big].astype(np.float64) - fp32[big]) / fp64[big]).max() if big.any() else 0:.3e})")
    assert rel_fp64.max() <= 2.0 ** -24
    assert rel_fp32.max() <= 1e-5
    assert count[nz].max() >= 300  # school-sized locations were reached


def test_fp32_order_at_the_largest_location_sizes():
    """Workplaces reach 2000 agents in the 10M-agent population, more than a
    100k-agent run puts at one infectious location: 200 locations each of
    500, 1000 and 2000 agents, every one infectious with a value of the
    shipped state and variant tables, summed in the reference's FP32
    ascending order, stay within 1e-5 of the exact sum that the engine
    rounds once (within 2^-24 of it)."""
    vals = np.array([np.float32(a) * np.float32(b) for a in (1.0, 0.75, 0.1)
                     for b in (1.0, 1.03, 1.79, 2.45, 2.7, 3.6)], dtype=np.float32)
    rng = np.random.default_rng(0)
    worst = 0.0
    for n in (500, 1000, 2000):
        x = rng.choice(vals, (200, n)).astype(np.float32)
        s = np.zeros(200, dtype=np.float32)
        for k in range(n):  # sequential float additions, one agent after the other
            s = (s + x[:, k]).astype(np.float32)
        exact = x.astype(np.float64).sum(axis=1)
        worst = max(worst, float(np.max(np.abs(s.astype(np.float64) - exact.astype(np.float32)) / exact)))
    print(f"max relative difference, reference FP32 order vs the engine's rounded exact sum: {worst:.3e}")
    assert worst <= 1e-5
