"""Calibration harness: the expected cumulative number of infections of a
simulation (or of an ensemble of replicas) as a differentiable function of
the infection coefficient k and the per-location-type infectiousness
multipliers m (m = 1 reproduces the reference).

The forward pass runs the GPU simulation with hazard accumulation on; the
gradient, with respect to k and m along the simulated trajectory (held
fixed), is accumulated by the infection kernel in the same pass
(pansim_hazard_* of include/pansim_hip.h), so the backward pass only scales
it by the incoming gradient.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
import torch

import pansim_amd

N_TYPES = 256  # multipliers are indexed by location-type id


def _steps_per_day(args):
    """Steps per simulated day of a command line (-t / --deltat minutes, 10 by default)."""
    a = [str(x) for x in args]
    for j, x in enumerate(a):
        if x in ("-t", "--deltat") and j + 1 < len(a):
            return 1440 // int(a[j + 1])
        if x.startswith("--deltat="):
            return 1440 // int(x.split("=", 1)[1])
    return 144


def _with_k(args: Sequence[str], k: float) -> List[str]:
    out, skip = [], False
    for a in args:
        if skip:
            skip = False
            continue
        if a in ("-k", "--infectionCoefficient"):
            skip = True
            continue
        out.append(str(a))
    return out + ["-k", repr(float(k))]


def run_hazard(replica_args: Sequence[Sequence[str]], k: float, m: np.ndarray, steps: int):
    """Runs the replicas (one batch) for `steps` steps and returns the mean
    over replicas of (J, dJ/dk, dJ/dm[256])."""
    m = np.ascontiguousarray(m, dtype=np.float32)
    reps = [_with_k(a, k) for a in replica_args]
    if len(reps) == 1:
        sims = [pansim_amd.Simulation(reps[0])]
        views = sims
    else:
        ens = pansim_amd.Ensemble(reps)
        sims = [ens]
        views = [ens.replica(r) for r in range(len(reps))]
    try:
        for v in views:
            v.set_location_type_multipliers(m)
            v.hazard_enable(True)
        spd = _steps_per_day(reps[0])
        if len(views) > 1 and steps % spd == 0:
            sims[0].days(steps // spd)  # one replay of the batch's day graph per simulated day
        else:
            sims[0].steps(steps)
        J, dk, dm = 0.0, 0.0, np.zeros(N_TYPES)
        for v in views:
            j, a, b = v.hazard()
            J, dk, dm = J + j, dk + a, dm + b
        n = len(views)
        return J / n, dk / n, dm / n
    finally:
        for s in sims:
            s.close()


class ExpectedInfections(torch.autograd.Function):
    """J(k, m): mean expected cumulative infections over the replicas.

    forward(k: 0-d tensor, m: 1-d tensor of multipliers by location-type id
    (shorter than 256: the rest are 1), replica_args, steps)."""

    @staticmethod
    def forward(ctx, k, m, replica_args, steps):
        mm = np.ones(N_TYPES, dtype=np.float32)
        mv = m.detach().cpu().double().numpy()
        mm[: len(mv)] = mv
        J, dk, dm = run_hazard(replica_args, float(k), mm, int(steps))
        ctx.save_for_backward(torch.tensor(dk, dtype=torch.float64), torch.tensor(dm[: len(mv)], dtype=torch.float64))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad):
        dk, dm = ctx.saved_tensors
        return grad * dk, grad * dm, None, None


def expected_infections(k: torch.Tensor, m: torch.Tensor, replica_args, steps: int) -> torch.Tensor:
    return ExpectedInfections.apply(k, m, replica_args, steps)
