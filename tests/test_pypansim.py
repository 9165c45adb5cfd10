"""The pybind11 module pyPanSim (the reference's Python binding,
matlab/mexPanSim.cpp:57-127, 207-219) and the ctypes mirror
pansim_amd.SimulatorInterface: the surface and the behaviour that needs no
device (CPU).  A run through the module is compared with the oracle in
test_gpu_pypansim.py."""
import subprocess
import sys

import pansim_amd

MSG = "Cannot run uninitialized simulation... Initialize it first."


def _run_py(code):
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd="/", env={"PYTHONPATH": pansim_amd.__path__[0] + "/..", "PATH": "/usr/bin:/bin"})
    return r


def test_module_surface():
    m = pansim_amd.pyPanSim()
    assert m.__name__ == "pyPanSim"
    assert m.__file__.startswith(pansim_amd.__path__[0])
    s = m.SimulatorInterface()
    for name in ("initSimulation", "runForDay", "close"):
        assert callable(getattr(s, name))


def test_uninitialized_run_for_day_prints_and_returns_empty():
    # printed from C, so checked in a child process's stdout
    r = _run_py("import pansim_amd; s = pansim_amd.pyPanSim().SimulatorInterface(); "
                "print(repr(s.runForDay(['TPdef'])), repr(s.runForDay([])))")
    assert r.returncode == 0, r.stderr
    assert r.stdout == MSG + MSG + "[] []\n"


def test_ctypes_mirror_uninitialized(capsys):
    s = pansim_amd.SimulatorInterface()
    assert s.runForDay(["TPdef"]) == []
    assert capsys.readouterr().out == MSG


def test_import_by_reference_name():
    # `import pyPanSim` with the module's directory on sys.path, as the
    # reference's scripts do with its build directory
    lib = pansim_amd.__path__[0] + "/_lib"
    r = _run_py(f"import sys; sys.path.insert(0, {lib!r}); import pyPanSim; print(pyPanSim.SimulatorInterface)")
    assert r.returncode == 0, r.stderr
    assert "pyPanSim.SimulatorInterface" in r.stdout
