"""Gradient of the expected number of infections with respect to k and the
location-type multipliers, along a trajectory held fixed.

CPU: the oracle's analytic FP64 gradient against central finite differences
of the same hazard sum (h = 1e-3 relative), evaluated along the same
trajectory; and m = 1 multipliers reproduce the plain run.
GPU: the device gradient (FP64 accumulation of per-agent terms whose
exponent is computed in f32) against the oracle's FP64 gradient, relative
tolerance 1e-5 (the workload definition asks for 1e-3)."""
from __future__ import annotations

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

TYPES = [2, 3, 4, 5, 6, 8, 12, 14, 22, 33, 44, 1]  # location-type ids of locationTypes.json


def _args(seed=1):
    return pansim_amd.config1_args(4000, 800, 1, seed, extra=["-k", "0.01"])


def _mult():
    m = np.ones(256)
    for j, t in enumerate(TYPES):
        m[t] = 0.8 + 0.05 * j
    return m


def test_oracle_gradient_matches_central_differences():
    o = OracleSim(_args())
    try:
        m = _mult()
        o.set_type_multipliers(m)
        o.hazard_enable(True)
        k = 0.01
        h = 1e-3
        probes = [(h * k, None), (-h * k, None)]
        for t in TYPES[:6]:
            d = np.zeros(256)
            d[t] = h * m[t]
            probes += [(0.0, d), (0.0, -d)]
        for dk, dm in probes:
            o.hazard_probe(dk, dm)
        o.steps(7 * 144)
        J, dk, dm, pj = o.hazard()
        assert J > 1.0
        fd_k = (pj[0] - pj[1]) / (2 * h * k)
        assert abs(fd_k - dk) <= 1e-3 * abs(dk)
        for j, t in enumerate(TYPES[:6]):
            fd = (pj[2 + 2 * j] - pj[3 + 2 * j]) / (2 * h * m[t])
            assert abs(fd - dm[t]) <= 1e-3 * max(abs(dm[t]), 1e-12), (t, fd, dm[t])
        # homogeneity: J depends on k and m only through k * m_type, so
        # k dJ/dk = sum_t m_t dJ/dm_t (with k and m as the simulation holds them, f32)
        k32, m32 = float(np.float32(k)), m.astype(np.float32).astype(np.float64)
        assert abs(k32 * dk - float(np.dot(m32, dm))) <= 1e-12 * abs(k32 * dk)
    finally:
        o.close()


def test_unit_multipliers_reproduce_the_plain_run():
    a, b = OracleSim(_args(3)), OracleSim(_args(3))
    try:
        b.set_type_multipliers(np.ones(256))
        a.steps(400)
        b.steps(400)
        for f in ("state", "location", "infectedTimestamp"):
            assert np.array_equal(a.field(f), b.field(f))
    finally:
        a.close()
        b.close()


@pytest.mark.gpu
def test_device_gradient_matches_oracle():
    args = _args(2)
    g = pansim_amd.Simulation(args)
    o = OracleSim(args)
    try:
        m = _mult()
        g.set_location_type_multipliers(m)
        o.set_type_multipliers(m)
        g.hazard_enable(True)
        o.hazard_enable(True)
        g.steps(7 * 144)
        o.steps(7 * 144)
        # the multipliers change the trajectory identically on both sides
        for f in ("state", "location", "infectedTimestamp", "susceptible0"):
            assert np.array_equal(g.field(f), o.field(f)), f
        Jg, dkg, dmg = g.hazard()
        Jo, dko, dmo, _ = o.hazard()
        assert Jo > 1.0
        assert abs(Jg - Jo) <= 1e-5 * abs(Jo)
        assert abs(dkg - dko) <= 1e-5 * abs(dko)
        assert np.all(np.abs(dmg - dmo) <= 1e-5 * np.maximum(np.abs(dmo), 1e-12))
    finally:
        g.close()
        o.close()


@pytest.mark.gpu
def test_autograd_function_matches_oracle():
    import torch
    from pansim_amd.calibration import expected_infections

    base = pansim_amd.config1_args(3000, 600, 1, 2)
    k = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    m = torch.ones(45, dtype=torch.float64)
    m[4], m[2] = 1.2, 0.9
    m.requires_grad_(True)
    J = expected_infections(k, m, [base], 5 * 144)
    J.backward()
    o = OracleSim(base + ["-k", repr(0.01)])
    try:
        mm = np.ones(256)
        mm[:45] = m.detach().numpy()
        o.set_type_multipliers(mm)
        o.hazard_enable(True)
        o.steps(5 * 144)
        Jo, dko, dmo, _ = o.hazard()
    finally:
        o.close()
    assert Jo > 1.0
    assert abs(J.item() - Jo) <= 1e-5 * Jo
    assert abs(k.grad.item() - dko) <= 1e-5 * abs(dko)
    assert np.all(np.abs(m.grad.numpy() - dmo[:45]) <= 1e-5 * np.maximum(np.abs(dmo[:45]), 1e-12))


@pytest.mark.gpu
def test_device_gradient_against_device_central_differences():
    """The reduced config-5 shape on the device: 32 batched replicas of
    20,000 agents (seed_i = i), 3 weeks, hazard accumulation with the
    finite-difference probes on: for every replica the device's dJ/dk and
    dJ/dm_work equal the central differences (J(theta(1+h)) - J(theta(1-h))) /
    (2 h theta) of J along the same trajectory (h = 1e-3, common random
    numbers: one run carries both) within 1e-3 relative."""
    h, work = 1e-3, 4
    reps = [pansim_amd.synthetic_args(20_000, 3, i, i + 1, extra=["-k", "0.002", "--synthInfected", "0.01"])
            for i in range(32)]
    ens = pansim_amd.Ensemble(reps)
    try:
        views = [ens.replica(r) for r in range(32)]
        for v in views:
            v.set_location_type_multipliers(np.ones(256, dtype=np.float32))
            v.hazard_enable(True)
            v.hazard_set_probe(h, work)
        ens.steps(3 * 7 * 144)
        k = np.float64(np.float32(0.002))
        worst = 0.0
        for v in views:
            out = v.hazard_with_probes()
            J, dk, dm = out[0], out[1], out[2:258]
            fd_k = (out[258] - out[259]) / (2 * h * k)
            fd_m = (out[260] - out[261]) / (2 * h * 1.0)
            assert J > 10.0
            assert abs(dk - fd_k) <= 1e-3 * abs(fd_k), (dk, fd_k)
            assert abs(dm[work] - fd_m) <= 1e-3 * abs(fd_m), (dm[work], fd_m)
            worst = max(worst, abs(dk - fd_k) / abs(fd_k), abs(dm[work] - fd_m) / abs(fd_m))
        print(f"largest relative difference gradient vs central differences: {worst:.2e}")
    finally:
        ens.close()
