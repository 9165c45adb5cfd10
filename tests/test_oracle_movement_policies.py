"""The oracle's DummyMovement and NoMovement (movementPolicies.h:9-65),
pinned by what the mathematics of the algorithm fixes (CPU): after the
midnight reset every agent draws a uniform location at the day's first
step (chi-square against the uniform law), and the number of moves of an
agent in one day follows the renewal process the countdown defines (waits
uniform on 0..35, a wait of 0 ending the agent's moves for the day through
the unsigned wrap), whose mean is computed exactly here; NoMovement never
changes a location."""
from __future__ import annotations

import numpy as np

import pansim_amd
from oracle.oracle import OracleSim


def _sim(policy, n=6000, nl=300, seed=5):
    return OracleSim(pansim_amd.with_data_dir(pansim_amd.config1_args(
        n, nl, 1, seed, extra=["--movementPolicy", policy])))


def expected_moves(steps=144, draws=36):
    """E[number of moves in `steps` steps] when a move happens at step 0 and
    each move draws w uniform on 0..draws-1: w = 0 stops, otherwise the next
    move happens w steps later."""
    e = np.zeros(steps + 1)  # e[r]: expected moves from a move with r steps left (r = steps - t)
    for r in range(1, steps + 1):
        # this move, then with w >= 1 another one w steps later if w < r
        e[r] = 1.0 + sum(e[r - w] for w in range(1, min(draws, r))) / draws
    return e[steps]


def test_dummy_first_step_uniform_and_renewal_mean():
    o = _sim("dummy")
    n, L = o.n_agents, o.n_locations
    o.steps(1)  # midnight (countdowns := 0), then everyone moves
    loc = o.field("location")
    counts = np.bincount(loc, minlength=L)
    chi2 = float(((counts - n / L) ** 2 / (n / L)).sum())
    # chi-square with L-1 degrees of freedom: mean L-1, sd sqrt(2(L-1))
    assert abs(chi2 - (L - 1)) < 6 * np.sqrt(2 * (L - 1))
    moves = np.ones(n)
    prev = loc
    for _ in range(143):
        o.steps(1)
        cur = o.field("location")
        moves += cur != prev  # a redraw of the same location (p = 1/L) is not seen
        prev = cur
    mean = expected_moves() * (1 - 1 / L) + 1 / L  # first move always counted
    sd = moves.std() / np.sqrt(n)
    assert abs(moves.mean() - mean) < 6 * sd + 0.02


def test_no_movement_keeps_every_location():
    o = _sim("none", n=3000, nl=200)
    start = o.field("location").copy()
    o.steps(2 * 144)
    assert np.array_equal(o.field("location"), start)


def test_no_testing_and_no_closure_policies():
    """NoTesting (testingPolicies.h:11-20): no agent is ever tested and the
    three testing entries of the statistics stay 0.  NoClosure
    (closurePolicies.h:36-46): no rule is read, so with the shipped closure
    rules (which close schools and workplaces from day 0 and put masks on)
    no location is ever closed and every infectiousness stays as loaded;
    with RuleClosure the same run does close locations."""
    base = pansim_amd.synthetic_args(3000, 1, 2, 2, extra=["-k", "0.01", "--synthInfected", "0.02"])
    o = OracleSim(pansim_amd.with_data_dir(base + ["--testingPolicy", "none"]))
    try:
        o.steps(7 * 144)
        assert (o.field("lastTest") == 0xFFFFFFFF).all() or (o.field("lastTest") == 0).all()
        for d in range(o.num_days):
            assert o.day_stats(d)[13:16] == [0, 0, 0]
    finally:
        o.close()
    ref = OracleSim(pansim_amd.with_data_dir(base))
    none = OracleSim(pansim_amd.with_data_dir(base + ["--closurePolicy", "none"]))
    try:
        inf0 = none.field("locInfectiousness").copy()
        closed_rule = 0
        for day in range(1, 7):
            ref.steps(day * 144 + 1 - ref.timestamp)
            none.steps(day * 144 + 1 - none.timestamp)
            assert (none.field("locOpen") == 1).all()
            assert np.array_equal(none.field("locInfectiousness"), inf0)
            closed_rule += int((ref.field("locOpen") == 0).sum())
        assert closed_rule > 0
    finally:
        ref.close()
        none.close()
