"""bench.py's contract on the CPU: more than one GPU is refused with a
message, and the committed HBM traffic of the movement kernel is attached
to a bench line only for the exact timed window it was measured over."""
import json
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_traffic_only_for_the_window_it_was_measured_over():
    with open(os.path.join(ROOT, "profiles", "traffic_k_move_tile_synth10M.json")) as fh:
        t = json.load(fh)
    assert bench.traffic_for(t["workload_key"]) == t["bytes_per_launch"]
    assert bench.traffic_for(t["workload_key"] + "x") is None
    assert bench.traffic_for("synth10000000_w4_s0_from0_k20") is None
    # the record adds up: scaled fetch + scaled write, per launch of the 20 window steps
    cal = t["copy_calibration"]
    kib = 1024.0
    fetch = sum(t["fetch_kib_per_launch_raw"]) / len(t["fetch_kib_per_launch_raw"]) * kib * cal["fetch_factor"]
    write = sum(t["write_kib_per_launch_raw"]) / len(t["write_kib_per_launch_raw"]) * kib * cal["write_factor"]
    assert len(t["fetch_kib_per_launch_raw"]) == len(t["write_kib_per_launch_raw"]) == 20
    assert abs(fetch + write - t["bytes_per_launch"]) <= 1e-6 * t["bytes_per_launch"]
    assert abs(cal["fetch_factor"] - cal["copy_bytes"] / (cal["fetch_kib_raw"] * kib)) < 1e-9
    assert abs(cal["write_factor"] - cal["copy_bytes"] / (cal["write_kib_raw"] * kib)) < 1e-9
