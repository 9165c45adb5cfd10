#!/usr/bin/env python3
"""Throughput of the plain GPU drop-in on its headline workload.

Workload: 10M-agent synthetic household population (schools, classrooms,
lognormal workplaces, shops, leisure, doctors, hospitals), 4 weeks at
dt = 10 min (4032 steps), k = 3.74e-4, 0.1% initially infected, all policies
on (DetailedTesting, quarantinePolicy 3, RuleClosure with the default
closureRules.json, non-COVID hospitalisation, immunity waning).

A "step" is one iteration of the reference's runSimulation loop: the
midnight block when due (testing, progression, immunity, other disease,
statistics, closures, location planning), then movement and infection for
every agent.  --warmup steps run untimed first (default: day 0), then exactly
--steps steps are timed between two device synchronisations (default: the
remaining 27 days).  value = agents * steps / seconds (agent-steps/s).

roofline: the per-step pipeline (movement kernel, then the infection kernel,
which also applies the movement's deferred location quarantine writes) is
the dominant work; its algorithmic bytes per step are
22 B/agent + 24 B/location + 4 B/relocated agent, divided by its device time
from HIP events on the context stream during the timed region: one pair per
day around the run of movement and infection launches between two midnight
blocks (inter-kernel gaps included, midnight work excluded).  The warmup
runs with an event pair per kernel launch instead, for the movement /
infection split.  traffic: HBM bytes per step of the
same kernels from rocprofv3 FETCH_SIZE / WRITE_SIZE passes
(profiles/traffic_synth10M.json, 10M-agent workload only), else null.

cpu_baseline: the oracle (sequential C++ re-statement of the reference
per-step pipeline) on a bounded sample of the same workload (1M synthetic
agents, one simulated week: 1008 steps), on rank 0 only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3888)
    ap.add_argument("--warmup", type=int, default=144)
    ap.add_argument("--agents", type=int, default=10_000_000)
    ap.add_argument("--weeks", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-agents", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=1008)
    ap.add_argument("--ensemble-replicas", type=int, default=64,
                    help="replicas of the batched-ensemble measurement (0 skips it)")
    ap.add_argument("--ensemble-steps", type=int, default=288)
    ap.add_argument("--gradient-replicas", type=int, default=32,
                    help="replicas of the gradient measurement (0 skips it)")
    ap.add_argument("--gradient-agents", type=int, default=1_000_000)
    ap.add_argument("--gradient-weeks", type=int, default=3)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (calibrates HBM byte counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain device-to-device copy of known size, for scaling the
    FETCH_SIZE / WRITE_SIZE counters of the step kernels."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4, "repeats": 5}), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(a):
    from oracle.oracle import OracleSim
    import pansim_amd

    args = pansim_amd.synthetic_args(a.cpu_agents, 1, a.seed, 1)
    o = OracleSim(args)
    t0 = time.perf_counter()
    o.steps(a.cpu_steps)
    dt = time.perf_counter() - t0
    return {
        "value": o.n_agents * a.cpu_steps / dt,
        "unit": "agent-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle (sequential C++ re-statement of the reference pipeline), {o.n_agents} synthetic agents, "
                  f"{a.cpu_steps} steps incl. the day-0 midnight block, {dt:.1f} s",
    }


def ensemble_args(n_rep, weeks):
    """Replica command lines of the batched-ensemble workload: sizes
    log-uniform in [1e5, 2e6], seed_i = i, k_i uniform in [2e-4, 6e-4],
    closure rules on for even replicas, quarantine policy i % 4."""
    import numpy as np
    import pansim_amd

    rng = np.random.default_rng(2024)
    sizes = np.exp(rng.uniform(np.log(1e5), np.log(2e6), n_rep)).astype(int)
    ks = rng.uniform(2e-4, 6e-4, n_rep)
    out = []
    for i in range(n_rep):
        extra = ["-k", f"{ks[i]:.6g}", "--quarantinePolicy", str(i % 4)]
        if i % 2:
            extra += ["--enableClosures", "0", "--closures", "inputConfigFiles/noRules.json"]
        out.append(pansim_amd.synthetic_args(int(sizes[i]), weeks, i, i + 1, extra=extra))
    return out, int(sizes.sum())


def ensemble_measure(a):
    """The same replicas stepped as one batch (one launch per kernel per step
    for all of them) and one after the other with the plain path; day 0 is
    warmup, then --ensemble-steps steps are timed in each mode."""
    import pansim_amd

    reps, total = ensemble_args(a.ensemble_replicas, a.weeks)
    K = a.ensemble_steps
    t0 = time.perf_counter()
    ens = pansim_amd.Ensemble(reps)
    log(f"ensemble: {a.ensemble_replicas} replicas, {total} agents built in {time.perf_counter() - t0:.1f}s")
    ens.steps(144)
    ens.sync()
    t1 = time.perf_counter()
    ens.steps(K)
    ens.sync()
    batched = time.perf_counter() - t1
    ens.close()
    log(f"ensemble: batched {K} steps in {batched:.3f}s")
    looped = 0.0
    for args in reps:
        sim = pansim_amd.Simulation(args)
        sim.steps(144)
        sim.sync()
        t1 = time.perf_counter()
        sim.steps(K)
        sim.sync()
        looped += time.perf_counter() - t1
        sim.close()
    log(f"ensemble: looped {K} steps in {looped:.3f}s")
    return {
        "workload": f"{a.ensemble_replicas} synthetic replicas, sizes log-uniform in [1e5, 2e6] (sum {total} agents), "
                    "seed_i = i, k_i ~ U[2e-4, 6e-4], closures on/off, quarantine policy i % 4, dt = 10 min",
        "replicas": a.ensemble_replicas,
        "agents": total,
        "steps": K,
        "warmup": 144,
        "batched_agent_steps_per_s": total * K / batched,
        "looped_agent_steps_per_s": total * K / looped,
        "speedup": looped / batched,
    }


def gradient_measure(a):
    """One evaluation of the expected cumulative infections and their
    gradient with respect to k and the location-type multipliers: the
    replicas (seed_i = i) run as one batch for the whole horizon with hazard
    accumulation on; the populations are built before the timed region."""
    import numpy as np
    import pansim_amd

    reps = [pansim_amd.synthetic_args(a.gradient_agents, a.gradient_weeks, i, i + 1) for i in range(a.gradient_replicas)]
    t0 = time.perf_counter()
    ens = pansim_amd.Ensemble(reps)
    log(f"gradient: {a.gradient_replicas} x {a.gradient_agents} agents built in {time.perf_counter() - t0:.1f}s")
    views = [ens.replica(r) for r in range(a.gradient_replicas)]
    m = np.ones(256, dtype=np.float32)
    steps = a.gradient_weeks * 7 * 144
    ens.sync()
    t1 = time.perf_counter()
    for v in views:
        v.set_location_type_multipliers(m)
        v.hazard_enable(True)
    ens.steps(steps)
    ens.sync()
    J = dk = 0.0
    for v in views:
        j, g, _ = v.hazard()
        J, dk = J + j, dk + g
    dt = time.perf_counter() - t1
    ens.close()
    log(f"gradient: one evaluation ({steps} steps) in {dt:.2f}s")
    return {
        "workload": f"{a.gradient_replicas} synthetic replicas x {a.gradient_agents} agents, {a.gradient_weeks} weeks, "
                    "gradient of the expected cumulative infections w.r.t. k and the location-type multipliers",
        "gradient_evals_per_s": 1.0 / dt,
        "agent_steps_per_s": a.gradient_replicas * a.gradient_agents * steps / dt,
        "mean_expected_infections": J / a.gradient_replicas,
        "mean_dJ_dk": dk / a.gradient_replicas,
    }


def main():
    a = parse()
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported (replicas would "
                                   "shard whole populations across GPUs with no collectives)"}))
        sys.exit(2)
    import pansim_amd

    total = a.weeks * 7 * 144
    if a.warmup + a.steps > total:
        a.steps = max(1, total - a.warmup)
    args = pansim_amd.synthetic_args(a.agents, a.weeks, a.seed, 1)
    t0 = time.perf_counter()
    sim = pansim_amd.Simulation(args)
    log(f"built {sim.n_agents} agents, {sim.n_locations} locations in {time.perf_counter() - t0:.1f}s")
    # the warmup runs with one event pair per kernel launch, for the split of
    # the step between movement and infection (informational)
    sim.enable_timing(True)
    sim.steps(a.warmup)
    sim.sync()
    wm, wmove_ms, wi, winfect_ms = sim.timing_summary()
    log(f"warmup {a.warmup} steps done (t={sim.timestamp})")
    sim.movement_counters(reset=True)
    # timed region: one event pair per day around its movement + infection launches
    sim.enable_timing(True, pipeline=True)
    sim.sync()
    t1 = time.perf_counter()
    done = 0
    chunk = 144
    while done < a.steps:
        n = min(chunk, a.steps - done)
        sim.steps(n)
        done += n
        if done % (7 * 144) == 0:
            log(f"  {done}/{a.steps} steps")
    sim.sync()
    elapsed = time.perf_counter() - t1
    n_pipe, pipe_ms, _, _ = sim.timing_summary()
    evaluated, relocated = sim.movement_counters()
    N, L = sim.n_agents, sim.n_locations
    K = a.steps
    value = N * K / elapsed
    # algorithmic bytes of the per-step pipeline (movement + infection)
    step_bytes_total = (22.0 * N + 24.0 * L) * K + 4.0 * relocated
    pipeline_s = pipe_ms / 1e3
    achieved = step_bytes_total / pipeline_s / 1e9 if pipeline_s > 0 else 0.0
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_synth10M.json")
    if N == 10_000_000 and os.path.exists(tpath):
        # HBM bytes per step of the same kernels, from the rocprofv3 FETCH_SIZE /
        # WRITE_SIZE passes recorded in that file (copy-calibrated)
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_step")
    last = sim.day_stats(sim.num_days - 1)
    sim.close()
    ensemble = None
    if a.ensemble_replicas > 0:
        ensemble = ensemble_measure(a)
    gradient = None
    if a.gradient_replicas > 0:
        gradient = gradient_measure(a)
    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(a)
    line = {
        "metric": "agent-steps/s",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+int64",
        "data": "synthetic",
        "config": {
            "workload": f"plain GPU drop-in, {N}-agent synthetic household population, {a.weeks} weeks at dt=10 min, "
                        "k=3.74e-4, 0.1% infected at start, all policies on",
            "n_agents": N,
            "n_locations": L,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "step_pipeline": {
            "pipeline_ms_per_step": pipe_ms / max(1, n_pipe),
            "midnight_and_gaps_ms_per_step": (elapsed * 1e3 - pipe_ms) / K,
            "warmup_movement_ms_per_step": wmove_ms / max(1, wm),
            "warmup_infection_ms_per_step": winfect_ms / max(1, wi),
            "movers_evaluated_per_step": evaluated / K,
            "relocated_per_step": relocated / K,
            "algorithmic_bytes_per_step": step_bytes_total / K,
        },
        "final_day_stats": last[:12],
        "ensemble": ensemble,
        "gradient": gradient,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
