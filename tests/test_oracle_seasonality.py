"""The oracle's seasonality multiplier (seasonalityMultiplier,
infectionPolicies.h:443-478) pinned by properties of the day arithmetic and
by values at points where the cosine is known exactly, instead of a second
copy of the formula:

- only startDate + day matters inside one regime, with period 366 days, and
  the day of the year is folded (x and 366 - x give the same value);
- the multiplier is 1 before the peak day (81 after the +6 offset) and at it;
- with c0 = 1 (the regime of days 251-449 takes c0 from --c0), 61 days past
  the peak the cosine is cos(pi/3) and the value is 0.5 * 0.5 + 0.5 = 0.75;
  with the default c0 = 3.08 that point clamps to the lower bound 0.5, and
  after day 600 the lower bound is --trunc;
- the regimes switch after day 250, before day 450 and after day 600, and
  --d_offset shifts the day inside the second and third regimes.
"""
import pytest

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144  # steps per day at dt = 10


def _sim(*extra):
    return OracleSim(pansim_amd.with_data_dir(pansim_amd.config1_args(50, 10, 1, 0, extra=list(extra))))


def _season(o, day):
    return o.seasonality(day * SPD)


@pytest.fixture(scope="module")
def sims():
    cache = {}

    def get(*extra):
        if extra not in cache:
            cache[extra] = _sim(*extra)
        return cache[extra]

    yield get
    for o in cache.values():
        o.close()


def test_only_the_date_matters_and_period(sims):
    base = sims("--startDate", "0")
    for k in (0, 17, 75, 120, 200, 250):
        shifted = sims("--startDate", str(k))
        assert _season(base, k) == _season(shifted, 0)
        assert _season(base, k) == _season(sims("--startDate", str(366)), k)


def test_fold_symmetry(sims):
    # day of the year x = startDate + day + 6; x and 366 - x give the same value
    for x in (90, 120, 150, 170, 183):
        a = sims("--startDate", str(x - 6))
        b = sims("--startDate", str(366 - x - 6))
        assert _season(a, 0) == _season(b, 0), x


def test_plateau_before_and_at_the_peak(sims):
    o = sims("--startDate", "0")
    for day in range(0, 81 - 6 + 1):  # x = 6 .. 81
        assert _season(o, day) == 1.0
    assert _season(o, 81 - 6 + 1) < 1.0  # one day past the peak


def test_known_cosine_points(sims):
    # second regime (days 251-449): c0 from --c0; x = 142 is 61 days past the
    # peak, cos(2 pi 61 / 366) = cos(pi / 3) = 1/2
    day = 300
    start = 142 - 6 - day + 366  # x = 142 (mod 366)
    o = sims("--startDate", str(start), "--c0", "1", "--d_offset", "0,0", "--d_peak_offset", "0")
    assert abs(_season(o, day) - 0.75) < 1e-12
    # default c0 = 3.08 in the first regime: 0.5 * 3.08 * 0.5 + 1 - 1.54 < 0.5, clamped
    o1 = sims("--startDate", str(142 - 6))
    assert _season(o1, 0) == 0.5
    # after day 600 the lower bound is --trunc
    o3 = sims("--startDate", str((142 - 6 - 700) % 366), "--trunc", "0.3", "--d_offset", "0,0")
    assert _season(o3, 700) == pytest.approx(0.3, abs=1e-12)


def test_regime_boundaries_and_offsets(sims):
    # a d_offset of 61 days moves x = 81 (the peak, value 1) to 142 inside the
    # second regime only: with c0 = 1 the value there is 0.75
    for day, second in ((250, False), (251, True), (449, True), (450, False)):
        start = (81 - 6 - day) % 366
        o = sims("--startDate", str(start), "--c0", "1", "--d_offset", "61,0", "--d_peak_offset", "0")
        v = _season(o, day)
        assert (abs(v - 0.75) < 1e-12) if second else v == 1.0, (day, v)
    # the third regime (after day 600) takes d_offset[1]
    for day, third in ((600, False), (601, True)):
        start = (81 - 6 - day) % 366
        o = sims("--startDate", str(start), "--d_offset", "0,61", "--trunc", "0.1")
        v = _season(o, day)
        # x = 142 with c0 = 3.08 gives 0.23 > 0.1: unclamped in the third regime
        assert (abs(v - (0.5 * 3.08 * 0.5 + 1.0 - 0.5 * 3.08)) < 1e-12) if third else v == 1.0, (day, v)
