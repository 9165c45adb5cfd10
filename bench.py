#!/usr/bin/env python3
"""Throughput of the plain GPU drop-in on its headline workload.

Workload: 10M-agent synthetic household population (schools, classrooms,
lognormal workplaces, shops, leisure, doctors, hospitals), 4 weeks at
dt = 10 min (4032 steps),
k = 3.74e-4, 0.1% initially infected, all policies on (DetailedTesting,
quarantinePolicy 3, RuleClosure with the default closureRules.json, non-COVID
hospitalisation, immunity waning).

A "step" is one iteration of the reference's runSimulation loop: the
midnight block when due (testing, progression, immunity, other disease,
statistics, closures, location planning), then movement and infection for
every agent.  value = agents * steps / seconds (agent-steps/s) over exactly
--steps timed steps between two device synchronisations.

Which steps are timed:
  * --steps >= 144 (default 3888 = days 1-27 after --warmup 144): the run
    from step --warmup on;
  * --steps < 144: a window inside the 4-week run, starting at
    --window-day / --window-time (default day 14, 07:00, the morning of the
    third week: the epidemic is under way and agents change activities).  The
    steps before it run untimed (spin-up, reported with its wall time), then
    --warmup untimed steps, then the window.  The window is named in
    config.workload.
A sub-record "full_day" times one complete simulated day (its midnight block
and 144 steps) after the timed steps, for the day-average picture.

roofline (the dominant kernel, k_move_tile, the movement step): per launch
the algorithmic bytes of the movement share of the per-step byte model,
8 B per agent (the reference's stepsUntilMove countdown, read and written)
+ 4 B per relocated agent (its new location), divided by the kernel's
average launch duration from a HIP event pair around every movement launch
of the timed steps (on the context stream).  traffic: HBM bytes per launch
of k_move_tile from rocprofv3 FETCH_SIZE / WRITE_SIZE passes over the same
window, copy-calibrated, read from profiles/traffic_<workload>.json when
that file names this exact workload and window, else null.
"step_pipeline" reports the movement + infection pipeline against the whole
per-step byte model 22 B/agent + 24 B/location + 4 B/relocated agent
(full_day record), and "movement_design" the bytes this design's movement
kernel itself must move (2 B/agent wake scan, 62 B per evaluated agent,
48 B per relocation).

cpu_baseline: the OpenMP build of the oracle (the reference's CPU structure:
parallel movement, a per-step stable sort by location, per-location sums in
ascending agent order) on the CPU-runnable workload at its size
(-n 100000 -N 20000, 2 weeks), on all host threads with OMP_PROC_BIND=true,
and with 1 thread (no sort, as the reference with one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STEPS_PER_DAY = 144


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3888)
    ap.add_argument("--warmup", type=int, default=144)
    ap.add_argument("--agents", type=int, default=10_000_000)
    ap.add_argument("--weeks", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--window-day", type=int, default=14)
    ap.add_argument("--window-time", default="07:00")
    ap.add_argument("--no-full-day", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="threads of the CPU baseline (0: all host threads)")
    ap.add_argument("--ensemble-replicas", type=int, default=64,
                    help="replicas of the batched-ensemble measurement (0 skips it)")
    ap.add_argument("--ensemble-days", type=int, default=7)
    ap.add_argument("--gradient-replicas", type=int, default=32,
                    help="replicas of the gradient measurement (0 skips it)")
    ap.add_argument("--gradient-agents", type=int, default=1_000_000)
    ap.add_argument("--gradient-weeks", type=int, default=3)
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (calibrates HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


CPU_SCRIPT = r"""
import sys, time, json
sys.path.insert(0, {root!r})
import pansim_amd
from oracle.oracle import OracleSim
o = OracleSim(pansim_amd.with_data_dir(pansim_amd.config1_args()), openmp=True)
t0 = time.perf_counter()
o.steps(2016)
dt = time.perf_counter() - t0
print(json.dumps({{"agents": o.n_agents, "steps": 2016, "seconds": dt}}))
"""


def cpu_run(threads):
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="true")
    r = subprocess.run([sys.executable, "-c", CPU_SCRIPT.format(root=ROOT)], env=env, capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed: " + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(a):
    # the host share of the GPU's box (OMP_NUM_THREADS there), else every thread
    threads = a.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    log(f"cpu baseline: config 1 on {threads} threads ...")
    many = cpu_run(threads)
    log(f"cpu baseline: {many['seconds']:.1f}s; 1 thread ...")
    one = cpu_run(1)
    rate = lambda r: r["agents"] * r["steps"] / r["seconds"]
    return {
        "value": rate(many),
        "unit": "agent-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"OpenMP oracle restatement of the reference's CPU path (parallel movement, per-step stable sort by "
                  f"location, per-location sums) on the CPU-runnable workload at its size: -n 100000 -N 20000 "
                  f"({many['agents']} agents after tourists are skipped, 20,001 locations), 2 weeks = 2016 steps, "
                  f"{threads} threads with OMP_PROC_BIND=true: {many['seconds']:.1f} s",
        "one_thread_value": rate(one),
        "one_thread_seconds": one["seconds"],
        "cpu_model": cpu_model(),
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
    """The same replicas stepped as one batch (every day one replay of the
    captured day graph: per step one movement and one infection launch for all
    replicas, the midnight blocks between the replays) and one after the
    other with the plain path; day 0 is warmup, then --ensemble-days whole
    days are timed in each mode."""
    import pansim_amd

    reps, total = ensemble_args(a.ensemble_replicas, a.weeks)
    D = a.ensemble_days
    K = D * 144
    t0 = time.perf_counter()
    ens = pansim_amd.Ensemble(reps)
    log(f"ensemble: {a.ensemble_replicas} replicas, {total} agents built in {time.perf_counter() - t0:.1f}s")
    ens.steps(144)
    ens.sync()
    t1 = time.perf_counter()
    ens.days(D)
    ens.sync()
    batched = time.perf_counter() - t1
    ens.close()
    log(f"ensemble: batched {D} days ({K} steps) in {batched:.3f}s")
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
        "mode": "batched: one hipGraph replay per simulated day; looped: the plain path per replica",
        "batched_agent_steps_per_s": total * K / batched,
        "looped_agent_steps_per_s": total * K / looped,
        "speedup": looped / batched,
    }


def gradient_measure(a):
    """One evaluation of the expected cumulative infections and their
    gradient with respect to k and the location-type multipliers: the
    replicas (seed_i = i) run as one batch for the whole horizon with hazard
    accumulation on, one hipGraph replay per simulated day; the populations
    are built before the timed region."""
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
    ens.days(a.gradient_weeks * 7)  # one replay of the day graph per simulated day
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


def clock(ts):
    m = (ts % STEPS_PER_DAY) * 10
    return f"day {ts // STEPS_PER_DAY} {m // 60:02d}:{m % 60:02d}"


def steps_timed(sim, n):
    """n steps between two synchronisations; returns the wall seconds."""
    sim.sync()
    t0 = time.perf_counter()
    done = 0
    while done < n:
        k = min(STEPS_PER_DAY, n - done)
        sim.steps(k)
        done += k
        if done % (7 * STEPS_PER_DAY) == 0:
            log(f"  {done}/{n} steps")
    sim.sync()
    dt = time.perf_counter() - t0
    return dt


def traffic_for(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic_k_move_tile_synth10M.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    if t.get("workload_key") != workload_key:
        return None
    return t.get("bytes_per_launch")


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

    total = a.weeks * 7 * STEPS_PER_DAY
    K, W = a.steps, a.warmup
    args = pansim_amd.synthetic_args(a.agents, a.weeks, a.seed, 1)
    t0 = time.perf_counter()
    sim = pansim_amd.Simulation(args)
    log(f"built {sim.n_agents} agents, {sim.n_locations} locations in {time.perf_counter() - t0:.1f}s")
    N, L = sim.n_agents, sim.n_locations
    window = K < STEPS_PER_DAY
    if window:
        hh, mm = (int(x) for x in a.window_time.split(":"))
        start = a.window_day * STEPS_PER_DAY + (hh * 60 + mm) // 10
        start = max(W, min(start, total - K - 2 * STEPS_PER_DAY))
        spin = start - W
    else:
        spin = 0
        K = min(K, total - W)
        start = W
    t1 = time.perf_counter()
    sim.steps(spin)
    sim.sync()
    spin_s = time.perf_counter() - t1
    if spin:
        log(f"spin-up: {spin} untimed steps to {clock(spin)} in {spin_s:.2f}s")
    sim.steps(W)
    sim.sync()
    log(f"warmup {W} steps done ({clock(sim.timestamp)})")

    # timed steps: one HIP event pair around every movement launch
    sim.movement_counters(reset=True)
    sim.enable_timing(True, movement_only=True)
    elapsed = steps_timed(sim, K)
    n_move, move_ms, _, _ = sim.timing_summary()
    sim.enable_timing(False)
    evaluated, relocated = sim.movement_counters()
    st = sim.day_stats(sim.num_days - 1)
    end = sim.timestamp
    window_text = f"steps {start}..{end - 1} ({clock(start)} to {clock(end - 1)})"
    log(f"timed {K} steps, {window_text}: {elapsed * 1e3 / K:.4f} ms/step")

    move_avg_s = move_ms / 1e3 / max(1, n_move)
    move_bytes = 8.0 * N + 4.0 * relocated / max(1, K)
    achieved = move_bytes / move_avg_s / 1e9 if move_avg_s > 0 else 0.0
    design_bytes = 2.0 * N + 62.0 * evaluated / max(1, K) + 48.0 * relocated / max(1, K)
    workload_key = f"synth{N}_w{a.weeks}_s{a.seed}_from{start}_k{K}"

    full_day = None
    if not a.no_full_day and sim.timestamp + 2 * STEPS_PER_DAY <= total:
        # the next complete simulated day: its midnight block and 144 steps
        nxt = (sim.timestamp // STEPS_PER_DAY + 1) * STEPS_PER_DAY
        sim.steps(nxt - sim.timestamp)
        sim.movement_counters(reset=True)
        sim.enable_timing(True, pipeline=True)
        day_s = steps_timed(sim, STEPS_PER_DAY)
        n_pipe, pipe_ms, _, _ = sim.timing_summary()
        sim.enable_timing(False)
        ev_d, rel_d = sim.movement_counters()
        b_step = 22.0 * N + 24.0 * L + 4.0 * rel_d / STEPS_PER_DAY
        pipe_s = pipe_ms / 1e3 / max(1, n_pipe)
        full_day = {
            "day": nxt // STEPS_PER_DAY,
            "agent_steps_per_s": N * STEPS_PER_DAY / day_s,
            "ms_per_step": day_s * 1e3 / STEPS_PER_DAY,
            "pipeline_ms_per_step": pipe_ms / max(1, n_pipe),
            "midnight_and_gaps_ms_per_step": (day_s * 1e3 - pipe_ms) / STEPS_PER_DAY,
            "movers_evaluated_per_step": ev_d / STEPS_PER_DAY,
            "relocated_per_step": rel_d / STEPS_PER_DAY,
            "step_pipeline": {
                "bytes_per_step": b_step,
                "achieved_GBs": b_step / pipe_s / 1e9 if pipe_s > 0 else 0.0,
                "frac": b_step / pipe_s / 1e9 / HBM_PEAK_GBS if pipe_s > 0 else 0.0,
                "model": "22 B/agent + 24 B/location + 4 B/relocated agent per step (movement + infection)",
            },
        }
        log(f"full day {full_day['day']}: {full_day['ms_per_step']:.4f} ms/step, pipeline "
            f"{full_day['pipeline_ms_per_step']:.4f} ms/step")
    sim.close()

    ensemble = ensemble_measure(a) if a.ensemble_replicas > 0 else None
    gradient = gradient_measure(a) if a.gradient_replicas > 0 else None
    cpu = cpu_baseline(a) if not a.no_cpu_baseline else None

    if window:
        workload = (f"plain GPU drop-in, {N}-agent synthetic household population, {a.weeks}-week run at dt=10 min, "
                    f"k=3.74e-4, 0.1% infected at start, all policies on; timed window {window_text} after "
                    f"{spin} untimed spin-up steps and {W} warmup steps")
    else:
        workload = (f"plain GPU drop-in, {N}-agent synthetic household population, {a.weeks} weeks at dt=10 min, "
                    f"k=3.74e-4, 0.1% infected at start, all policies on; timed {window_text} after {W} warmup steps")
    line = {
        "metric": "agent-steps/s",
        "value": N * K / elapsed,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32+fp32",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "n_agents": N,
            "n_locations": L,
            "window_first_step": start,
            "spinup_steps": spin,
            "spinup_seconds": spin_s,
            "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_move_tile",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_for(workload_key),
            "bytes_per_launch": move_bytes,
            "avg_launch_ms": move_avg_s * 1e3,
            "launches": n_move,
            "model": "8 B/agent (stepsUntilMove read + write) + 4 B/relocated agent, per movement launch",
            "workload_key": workload_key,
        },
        "movement_design": {
            "bytes_per_launch": design_bytes,
            "achieved_GBs": design_bytes / move_avg_s / 1e9 if move_avg_s > 0 else 0.0,
            "frac": design_bytes / move_avg_s / 1e9 / HBM_PEAK_GBS if move_avg_s > 0 else 0.0,
            "model": "2 B/agent wake scan + 62 B/evaluated agent + 48 B/relocation",
            "movers_evaluated_per_step": evaluated / max(1, K),
            "relocated_per_step": relocated / max(1, K),
        },
        "window_day_stats": st[:12],
        "full_day": full_day,
        "ensemble": ensemble,
        "gradient": gradient,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
