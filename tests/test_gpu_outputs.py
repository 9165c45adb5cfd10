"""GPU parity of the output files: the --outAgentStat JSON, the
--dumpLocationInfectiousList dump and the --diags summary of a HIP run are
byte-identical to the oracle's for the same arguments (the dump collects the
infections of every variant pass on the device and orders them on the host;
the oracle restates both writers independently)."""
import pytest

import pansim_amd
from oracle.oracle import OracleSim

pytestmark = pytest.mark.gpu


def _args(tmp, tag, extra=()):
    return pansim_amd.synthetic_args(3000, 1, 2, 4, extra=[
        "-k", "0.02", "--synthInfected", "0.03", "--outAgentStat", str(tmp / f"agents_{tag}.json"),
        "--dumpLocationInfectiousList", str(tmp / f"infectious_{tag}.txt"), "--diags", "1", *extra])


@pytest.mark.parametrize("steps", [3 * 144, 3 * 144 + 70])
def test_output_files_match_oracle(tmp_path, capfd, steps):
    sim = pansim_amd.Simulation(_args(tmp_path, "gpu"))
    sim.steps(steps)
    capfd.readouterr()
    sim.finalize()
    gpu_diags = capfd.readouterr().out
    sim.close()
    o = OracleSim(pansim_amd.with_data_dir(_args(tmp_path, "cpu")))
    o.steps(steps)
    cpu_diags = o.finalize()
    assert (tmp_path / "agents_gpu.json").read_text() == (tmp_path / "agents_cpu.json").read_text()
    gpu_dump = (tmp_path / "infectious_gpu.txt").read_text()
    assert gpu_dump == (tmp_path / "infectious_cpu.txt").read_text()
    assert int(gpu_dump.split("\n")[0]) > 0  # the run infected someone
    assert gpu_diags == cpu_diags


def test_output_files_through_run_for_day(tmp_path, capfd):
    # the day interface (SimulatorInterface): files written at close
    si = pansim_amd.SimulatorInterface()
    si.initSimulation(pansim_amd.with_data_dir(_args(tmp_path, "gpu")))
    for _ in range(3):
        si.runForDay(["TPdef"])
    capfd.readouterr()
    si.close()
    gpu_diags = capfd.readouterr().out
    o = OracleSim(pansim_amd.with_data_dir(_args(tmp_path, "cpu")))
    for _ in range(3):
        o.run_for_day(["TPdef"])
    cpu_diags = o.finalize()
    assert (tmp_path / "agents_gpu.json").read_text() == (tmp_path / "agents_cpu.json").read_text()
    assert (tmp_path / "infectious_gpu.txt").read_text() == (tmp_path / "infectious_cpu.txt").read_text()
    assert gpu_diags == cpu_diags
