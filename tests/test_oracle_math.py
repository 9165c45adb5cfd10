"""Pins the oracle's arithmetic building blocks against published known
answers and against what the mathematics fixes (CPU only).

- Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11).
- exp (f32) / log (f64): at most 1 ulp from correctly rounded values
  (float64 exp rounded to f32; 80-bit long double log rounded to f64).
- Truncated-geometric parameter (basicLengthAbstract.cpp:4-29): the chosen p
  gives the requested mean under the closed-form truncated geometric mean.
- Geometric draw (randomGenerator.h:55-63 shape): P(X=0) = p and
  E[X] = (1-p)/p on many independent Philox streams.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O


# Random123 KAT vectors for philox4x32_10 (kat_vectors file of Random123 1.09)
PHILOX_KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expected", PHILOX_KAT)
def test_philox_known_answers(ctr, key, expected):
    assert tuple(O.philox(ctr, key)) == expected


def test_philox_counter_sensitivity():
    # every input word changes the output (a dropped word in the round function fails here)
    base = O.philox((1, 2, 3, 4), (5, 6))
    for i in range(4):
        c = [1, 2, 3, 4]
        c[i] ^= 1
        assert O.philox(tuple(c), (5, 6)) != base
    assert O.philox((1, 2, 3, 4), (5 ^ 1, 6)) != base
    assert O.philox((1, 2, 3, 4), (5, 6 ^ 1)) != base


def _ulp32(x: np.ndarray) -> np.ndarray:
    x = np.abs(x.astype(np.float32))
    return np.spacing(x).astype(np.float64)


def test_expf_within_one_ulp():
    rng = np.random.default_rng(7)
    xs = np.concatenate([
        rng.uniform(-103.0, 88.7, 60000),
        -np.exp(rng.uniform(np.log(1e-12), np.log(100.0), 60000)),  # the infection ratio's range
        np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 88.0, -87.0, -100.0, 1e-30, -1e-30,
                  np.log(2.0), -np.log(2.0), 0.34657359, -0.34657359]),
    ]).astype(np.float32)
    got = np.array([O.expf(float(x)) for x in xs], dtype=np.float32)
    ref = np.exp(xs.astype(np.float64)).astype(np.float32)
    err = np.abs(got.astype(np.float64) - ref.astype(np.float64))
    tol = np.maximum(_ulp32(ref), np.finfo(np.float32).smallest_subnormal)
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, f"expf off by more than 1 ulp at x={xs[bad[:5]].tolist()}"
    assert O.expf(0.0) == 1.0
    assert O.expf(-200.0) == 0.0
    assert math.isinf(O.expf(100.0))
    assert math.isnan(O.expf(float("nan")))


def test_log_within_one_ulp():
    rng = np.random.default_rng(11)
    xs = np.concatenate([
        np.exp(rng.uniform(-700.0, 700.0, 40000)),
        1.0 + rng.uniform(-0.3, 0.3, 40000),
        1.0 - np.ldexp(1.0, -rng.integers(1, 53, 2000)),   # 1 - u: the geometric draw's argument
        np.array([0.5, 2.0, 1.0, 0.70710678118654757, 1.4142135623730951, 5e-324, 1e-310]),
    ])
    got = np.array([O.log(float(x)) for x in xs])
    ref = np.log(xs.astype(np.longdouble)).astype(np.float64)
    err = np.abs(got - ref)
    tol = np.spacing(np.abs(ref))
    tol[ref == 0.0] = 0.0
    bad = np.nonzero(err > tol)[0]
    assert bad.size == 0, f"log off by more than 1 ulp at x={xs[bad[:5]].tolist()}"
    assert O.log(1.0) == 0.0
    assert math.isinf(O.log(0.0)) and O.log(0.0) < 0
    assert math.isnan(O.log(-1.0))


def _truncated_geometric_mean(p: float, n: int) -> float:
    """Mean of (i+1) for i in [0, n) with weights (1-p)^i p, in closed form."""
    q = 1.0 - p
    if n <= 0:
        return float("nan")
    num = (1.0 - (n + 1) * q ** n + n * q ** (n + 1)) / (1.0 - q) ** 2
    den = (1.0 - q ** n) / (1.0 - q)
    return num / den


# (average, maximum) pairs of the shipped transition matrices' state lengths
@pytest.mark.parametrize("avg,mx", [(5.0, 14.0), (3.0, 7.0), (10.0, 28.0), (2.5, 6.0), (7.0, 21.0), (1.5, 4.0)])
def test_length_parameter_matches_truncated_mean(avg, mx):
    p = O.length_p(avg, mx)
    assert 0.0 < p <= 1.0
    assert abs(_truncated_geometric_mean(p, int(round(mx))) - avg) <= 1.1e-4


def test_length_parameter_special_cases():
    assert O.length_p(-1.0, 5.0) == 1.0
    # maximum -1 means "no maximum": float max, whose (unsigned) round is 0 on x86,
    # so the secant iteration stops at once at its second start point 0.9/avg
    fmax = float(np.finfo(np.float32).max)
    assert O.length_p(4.0, fmax) == pytest.approx(0.9 / 4.0, rel=0, abs=1e-15)


def test_geometric_distribution():
    p = 0.2
    draws = np.array([O.geometric(3, i, p) for i in range(40000)])
    assert draws.min() == 0
    n = len(draws)
    mean = (1 - p) / p
    sd = math.sqrt((1 - p) / p ** 2 / n)
    assert abs(draws.mean() - mean) < 5 * sd
    p0 = (draws == 0).mean()
    assert abs(p0 - p) < 5 * math.sqrt(p * (1 - p) / n)
