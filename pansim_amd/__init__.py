"""pansim_amd: MI355X-native PanSim per-timestep agent pipeline.

Python front end of the native library (pansim_amd/_lib/libpansim.so, built
by pansim_amd/build.py): the host Simulation<...> loop with HIP-backed
policies, reached through the C ABI of include/pansim_sim.h and
include/pansim_hip.h.  There is no Python or CPU fallback: if the library is
missing, or no GPU is present, the calls raise.

    sim = pansim_amd.Simulation(["-n", "100000", "-N", "20000", "-w", "2"])
    sim.run()                      # Simulation::runSimulation
    stats = sim.day_stats(3)       # per-day statistics vector

    si = pansim_amd.SimulatorInterface()   # pyPanSim.SimulatorInterface
    si.initSimulation([...]); si.runForDay(["TPdef", "QU3"])
"""
from __future__ import annotations

import ctypes
import os
import sys
from typing import List, Sequence

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "_lib")
DATA_DIR = os.path.join(PKG_DIR, "data")
LIB_PATH = os.path.join(LIB_DIR, "libpansim.so")

_lib = None

AGENT_FIELDS = [
    "state", "variant", "days", "location", "stepsUntilMove", "diagnosed", "quarantined", "stayedHome",
    "noWorkAgent", "lastTest", "susceptible0", "infectedTimestamp", "infectedLocation", "worstStateTimestamp",
    "worstStateEndTimestamp", "diagnosedTimestamp", "quarantinedTimestamp", "quarantinedUntilTimestamp",
    "daysInQuarantine", "hospitalizedTimestamp", "hospitalizedUntilTimestamp", "recoveredTimestamp",
    "infectedCount", "statsVariant", "worstState", "immunizationTimestamp", "immunizationCount", "vaccinationRound",
]
LOCATION_FIELDS = ["locQuarantineUntil", "locClosedUntil", "locOpen", "locInfectiousness", "locCount"]


class PanSimError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Loads libpansim.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"native library missing: {LIB_PATH} (run pansim_amd/build.py)")
    L = ctypes.CDLL(LIB_PATH)
    c_u32p = ctypes.POINTER(ctypes.c_uint32)
    L.pansim_sim_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int]
    L.pansim_sim_create_error.restype = ctypes.c_char_p
    L.pansim_sim_error.argtypes = [ctypes.c_void_p]
    L.pansim_sim_error.restype = ctypes.c_char_p
    L.pansim_sim_destroy.argtypes = [ctypes.c_void_p]
    L.pansim_sim_steps.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_sim_run.argtypes = [ctypes.c_void_p]
    L.pansim_sim_finalize.argtypes = [ctypes.c_void_p]
    L.pansim_sim_run_for_day.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_char_p), c_u32p,
                                         ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    for f in ("pansim_sim_timestamp", "pansim_sim_steps_per_day", "pansim_sim_total_steps"):
        getattr(L, f).argtypes = [ctypes.c_void_p]
        getattr(L, f).restype = ctypes.c_uint32
    L.pansim_sim_num_days.argtypes = [ctypes.c_void_p]
    L.pansim_sim_day_stats.argtypes = [ctypes.c_void_p, ctypes.c_int, c_u32p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_size_t)]
    L.pansim_sim_header.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    L.pansim_sim_sizes.argtypes = [ctypes.c_void_p, c_u32p, c_u32p, c_u32p]
    L.pansim_sim_context.argtypes = [ctypes.c_void_p]
    L.pansim_sim_context.restype = ctypes.c_void_p
    L.pansim_download_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, c_u32p, ctypes.c_size_t]
    L.pansim_download_location_sums.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int64),
                                                ctypes.c_size_t]
    L.pansim_last_error.argtypes = [ctypes.c_void_p]
    L.pansim_last_error.restype = ctypes.c_char_p
    L.pansim_last_kernel_ms.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.pansim_enable_timing.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_timing_summary.argtypes = [ctypes.c_void_p, c_u32p, ctypes.POINTER(ctypes.c_double), c_u32p,
                                        ctypes.POINTER(ctypes.c_double)]
    L.pansim_movement_counters.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32]
    L.pansim_sync.argtypes = [ctypes.c_void_p]
    L.pansim_device_expf.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]
    L.pansim_device_log.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_size_t]
    L.pansim_device_philox.argtypes = [c_u32p, ctypes.c_uint32, ctypes.c_uint32, c_u32p, ctypes.c_size_t]
    L.pansim_ensemble_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint32,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.POINTER(ctypes.c_char_p))]
    L.pansim_ensemble_create_error.restype = ctypes.c_char_p
    L.pansim_ensemble_error.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_error.restype = ctypes.c_char_p
    L.pansim_ensemble_destroy.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_steps.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_ensemble_days.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_ensemble_batch.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_batch.restype = ctypes.c_void_p
    L.pansim_batch_set_tiles.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
    L.pansim_batch_tiles.argtypes = [ctypes.c_void_p, c_u32p, c_u32p]
    L.pansim_batch_error.argtypes = [ctypes.c_void_p]
    L.pansim_batch_error.restype = ctypes.c_char_p
    L.pansim_ensemble_run.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_sync.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_size.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_size.restype = ctypes.c_uint32
    L.pansim_ensemble_timestamp.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_timestamp.restype = ctypes.c_uint32
    L.pansim_ensemble_steps_per_day.argtypes = [ctypes.c_void_p]
    L.pansim_ensemble_steps_per_day.restype = ctypes.c_uint32
    L.pansim_ensemble_context.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_ensemble_context.restype = ctypes.c_void_p
    L.pansim_ensemble_num_days.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_ensemble_day_stats.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, c_u32p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_size_t)]
    L.pansim_ensemble_sizes.argtypes = [ctypes.c_void_p, ctypes.c_uint32, c_u32p, c_u32p, c_u32p]
    L.pansim_set_location_type_multipliers.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_uint32]
    L.pansim_hazard_enable.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    L.pansim_hazard_read.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32, ctypes.c_uint32]
    L.pansim_hazard_set_probe.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_uint32]
    _lib = L
    return L


def _argv(args: Sequence[str]):
    arr = (ctypes.c_char_p * max(1, len(args)))()
    for i, a in enumerate(args):
        arr[i] = str(a).encode()
    return arr


def with_data_dir(args: Sequence[str]) -> List[str]:
    """Adds --dataDir pointing at the shipped input files unless given."""
    args = [str(a) for a in args]
    if "--dataDir" not in args:
        args += ["--dataDir", DATA_DIR]
    return args


class Simulation:
    """The reference's Simulation<...> (runSimulation / runForDay) on the GPU."""

    def __init__(self, args: Sequence[str], print_output: bool = False):
        L = lib()
        self._args = with_data_dir(args)
        h = ctypes.c_void_p()
        rc = L.pansim_sim_create(ctypes.byref(h), len(self._args), _argv(self._args), 1 if print_output else 0)
        if rc != 0:
            raise PanSimError(rc, L.pansim_sim_create_error().decode())
        self._h = h
        na, nl, nv = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        L.pansim_sim_sizes(self._h, ctypes.byref(na), ctypes.byref(nl), ctypes.byref(nv))
        self.n_agents, self.n_locations, self.n_variants = na.value, nl.value, nv.value

    def close(self):
        if getattr(self, "_h", None):
            lib().pansim_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PanSimError(rc, lib().pansim_sim_error(self._h).decode())

    @property
    def ctx(self):
        return lib().pansim_sim_context(self._h)

    def steps(self, n: int):
        self._check(lib().pansim_sim_steps(self._h, n))

    def run(self):
        """runSimulation: every step, then the output files."""
        self._check(lib().pansim_sim_run(self._h))

    def finalize(self):
        """The output files (--outAgentStat, --dumpLocationInfectiousList) and
        the --diags summary of a simulation driven by steps or run_for_day."""
        self._check(lib().pansim_sim_finalize(self._h))

    def run_for_day(self, flags: Sequence[str] = ()) -> List[int]:
        out = (ctypes.c_uint32 * 512)()
        n = ctypes.c_size_t()
        self._check(lib().pansim_sim_run_for_day(self._h, len(flags), _argv(flags), out, 512, ctypes.byref(n)))
        return list(out[: n.value])

    @property
    def timestamp(self) -> int:
        return lib().pansim_sim_timestamp(self._h)

    @property
    def steps_per_day(self) -> int:
        return lib().pansim_sim_steps_per_day(self._h)

    @property
    def total_steps(self) -> int:
        return lib().pansim_sim_total_steps(self._h)

    @property
    def num_days(self) -> int:
        return lib().pansim_sim_num_days(self._h)

    def day_stats(self, d: int) -> List[int]:
        out = (ctypes.c_uint32 * 512)()
        n = ctypes.c_size_t()
        self._check(lib().pansim_sim_day_stats(self._h, d, out, 512, ctypes.byref(n)))
        return list(out[: n.value])

    @property
    def header(self) -> List[str]:
        buf = ctypes.create_string_buffer(4096)
        lib().pansim_sim_header(self._h, buf, 4096)
        return [x for x in buf.value.decode().split("\t") if x]

    def field(self, name: str) -> np.ndarray:
        n = self.n_locations if name in LOCATION_FIELDS else self.n_agents
        out = np.zeros(n, dtype=np.uint32)
        rc = lib().pansim_download_field(self.ctx, name.encode(), self.timestamp,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), n)
        if rc != 0:
            raise PanSimError(rc, lib().pansim_last_error(self.ctx).decode())
        return out

    def possible_locations(self):
        """(first entry, number of entries) of every agent's possible-location
        list, and all list entries, as held on the device."""
        begin, count = self.field("csrBegin"), self.field("csrCount")
        total = int(begin[-1]) + int(count[-1]) if self.n_agents else 0
        out = np.zeros(max(1, total), dtype=np.uint32)
        rc = lib().pansim_download_field(self.ctx, b"possibleLocations", self.timestamp,
                                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), len(out))
        if rc != 0:
            raise PanSimError(rc, lib().pansim_last_error(self.ctx).decode())
        return begin, count, out[:total]

    def enable_timing(self, on: bool = True, pipeline: bool = False, movement_only: bool = False):
        """HIP-event timing of movement and infection: one event pair per call
        (kernel mode), only per movement call (movement_only=True) or, with
        pipeline=True, one pair around each run of movement + infection steps
        between other calls (a day's steps)."""
        mode = 2 if pipeline else 3 if movement_only else 1
        self._check_ctx(lib().pansim_enable_timing(self.ctx, mode if on else 0))

    def timing_summary(self):
        """(n_movement_calls, movement_ms, n_infection_calls, infection_ms) since enable_timing;
        in pipeline mode (n_steps, device ms of their movement + infection launches, 0, 0)."""
        nm, ni = ctypes.c_uint32(), ctypes.c_uint32()
        mm, mi = ctypes.c_double(), ctypes.c_double()
        self._check_ctx(lib().pansim_timing_summary(self.ctx, ctypes.byref(nm), ctypes.byref(mm), ctypes.byref(ni),
                                                    ctypes.byref(mi)))
        return nm.value, mm.value, ni.value, mi.value

    def movement_counters(self, reset: bool = False):
        """(agents whose movement routine ran, agents that changed location) since the last reset."""
        e, r = ctypes.c_uint64(), ctypes.c_uint64()
        self._check_ctx(lib().pansim_movement_counters(self.ctx, ctypes.byref(e), ctypes.byref(r), 1 if reset else 0))
        return e.value, r.value

    def sync(self):
        self._check_ctx(lib().pansim_sync(self.ctx))

    def _check_ctx(self, rc):
        if rc != 0:
            raise PanSimError(rc, lib().pansim_last_error(self.ctx).decode())

    def set_location_type_multipliers(self, m):
        """Infectiousness multiplier of every location type id (1 = reference)."""
        a = np.ascontiguousarray(m, dtype=np.float32)
        self._check_ctx(lib().pansim_set_location_type_multipliers(
            self.ctx, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(a)))

    def hazard_enable(self, on: bool = True):
        self._check_ctx(lib().pansim_hazard_enable(self.ctx, 1 if on else 0))

    def hazard(self, reset: bool = False):
        """(J, dJ/dk, dJ/dm[256]): expected infections and their gradient so far."""
        out = self.hazard_with_probes(reset)
        return out[0], out[1], out[2:258]

    def hazard_set_probe(self, h: float, loc_type: int):
        """Finite-difference probes along the trajectory: J at k(1 +- h) and at
        the multiplier of loc_type scaled by (1 +- h) (h = 0: off)."""
        self._check_ctx(lib().pansim_hazard_set_probe(self.ctx, h, loc_type))

    def hazard_with_probes(self, reset: bool = False) -> np.ndarray:
        """[J, dJ/dk, dJ/dm[256], J(k+) - J, J(k-) - J, J(m_p+) - J, J(m_p-) - J]."""
        out = np.zeros(262)
        self._check_ctx(lib().pansim_hazard_read(self.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 262,
                                                 1 if reset else 0))
        return out

    def location_sums(self, variant: int) -> np.ndarray:
        out = np.zeros(self.n_locations, dtype=np.int64)
        rc = lib().pansim_download_location_sums(self.ctx, variant, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                                 self.n_locations)
        if rc != 0:
            raise PanSimError(rc, lib().pansim_last_error(self.ctx).decode())
        return out


class Ensemble:
    """Replicas (each from its own command line: population, seed, k,
    policies) advanced in lockstep with one movement and one infection launch
    per step for all of them.  replica(r) is a read-only view with the
    Simulation accessors (fields, location sums, per-day statistics)."""

    def __init__(self, replica_args: Sequence[Sequence[str]]):
        L = lib()
        self._args = [with_data_dir(a) for a in replica_args]
        n = len(self._args)
        argcs = (ctypes.c_int * n)(*[len(a) for a in self._args])
        self._argvs = [_argv(a) for a in self._args]
        argvs = (ctypes.POINTER(ctypes.c_char_p) * n)(*[ctypes.cast(a, ctypes.POINTER(ctypes.c_char_p)) for a in self._argvs])
        h = ctypes.c_void_p()
        rc = L.pansim_ensemble_create(ctypes.byref(h), n, argcs, argvs)
        if rc != 0:
            raise PanSimError(rc, L.pansim_ensemble_create_error().decode())
        self._h = h
        self.size = n

    def close(self):
        if getattr(self, "_h", None):
            lib().pansim_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise PanSimError(rc, lib().pansim_ensemble_error(self._h).decode())

    def steps(self, n: int):
        self._check(lib().pansim_ensemble_steps(self._h, n))

    def set_tiles(self, move_items: int = 0, infect_items: int = 0):
        """Fixes the batch kernels' agents per thread (movement 8/16/32,
        infection 4/8), or (0, 0): let the autotuner pick them."""
        b = lib().pansim_ensemble_batch(self._h)
        if lib().pansim_batch_set_tiles(b, move_items, infect_items) != 0:
            raise PanSimError(-1, lib().pansim_batch_error(b).decode())

    def tiles(self):
        """(movement, infection) agents per thread in use; (0, 0) while tuning."""
        m, i = ctypes.c_uint32(), ctypes.c_uint32()
        lib().pansim_batch_tiles(lib().pansim_ensemble_batch(self._h), ctypes.byref(m), ctypes.byref(i))
        return m.value, i.value

    def days(self, n: int):
        """n whole days from a midnight, each one replay of the captured day graph."""
        self._check(lib().pansim_ensemble_days(self._h, n))

    def run(self):
        self._check(lib().pansim_ensemble_run(self._h))

    def sync(self):
        self._check(lib().pansim_ensemble_sync(self._h))

    @property
    def timestamp(self) -> int:
        return lib().pansim_ensemble_timestamp(self._h)

    def replica(self, r: int) -> "EnsembleReplica":
        return EnsembleReplica(self, r)


class EnsembleReplica:
    """Read-only view of one replica of an Ensemble."""

    def __init__(self, ens: Ensemble, r: int):
        if not 0 <= r < ens.size:
            raise IndexError(r)
        self._e, self._r = ens, r
        na, nl, nv = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        lib().pansim_ensemble_sizes(ens._h, r, ctypes.byref(na), ctypes.byref(nl), ctypes.byref(nv))
        self.n_agents, self.n_locations, self.n_variants = na.value, nl.value, nv.value
        self.steps_per_day = lib().pansim_ensemble_steps_per_day(ens._h)

    @property
    def ctx(self):
        return lib().pansim_ensemble_context(self._e._h, self._r)

    @property
    def timestamp(self) -> int:
        return self._e.timestamp

    @property
    def num_days(self) -> int:
        return lib().pansim_ensemble_num_days(self._e._h, self._r)

    def day_stats(self, d: int) -> List[int]:
        out = (ctypes.c_uint32 * 512)()
        n = ctypes.c_size_t()
        rc = lib().pansim_ensemble_day_stats(self._e._h, self._r, d, out, 512, ctypes.byref(n))
        if rc != 0:
            raise PanSimError(rc, "day_stats")
        return list(out[: n.value])

    field = Simulation.field
    location_sums = Simulation.location_sums
    set_location_type_multipliers = Simulation.set_location_type_multipliers
    hazard_enable = Simulation.hazard_enable
    hazard = Simulation.hazard
    hazard_set_probe = Simulation.hazard_set_probe
    hazard_with_probes = Simulation.hazard_with_probes
    _check_ctx = Simulation._check_ctx


class SimulatorInterface:
    """Mirror of the reference's pybind11 pyPanSim.SimulatorInterface
    (matlab/mexPanSim.cpp:57-127, 207-219); the compiled module of that name
    is pyPanSim().  As there, runForDay with an empty flag list runs nothing
    and returns the previous day's statistics (mexPanSim.cpp:113-114), and
    runForDay before initSimulation prints a message and returns []."""

    def __init__(self):
        self.sim = None
        self._stats: List[int] = []

    def initSimulation(self, args: Sequence[str]) -> None:
        args = list(args)
        if args and not str(args[0]).startswith("-"):
            args = args[1:]  # the reference receives argv including the program name
        self.close()
        self.sim = Simulation(args)
        self._stats = []

    def runForDay(self, flags: Sequence[str]) -> List[int]:
        if self.sim is None:
            print("Cannot run uninitialized simulation... Initialize it first.", end="", flush=True)
            return []
        flags = list(flags)
        if not flags:
            return self._stats
        self._stats = self.sim.run_for_day(flags)
        return self._stats

    def close(self) -> None:
        """The reference's destructor: finalize (output files), then release."""
        if self.sim is not None:
            try:
                self.sim.finalize()
            finally:
                self.sim.close()
                self.sim = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pyPanSim():
    """The pybind11 module pyPanSim (the reference's binding name, built
    next to libpansim.so): ``import pyPanSim`` works once
    ``pansim_amd/_lib`` is on sys.path; this returns the module either way."""
    lib = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
    if lib not in sys.path:
        sys.path.insert(0, lib)
    import importlib
    return importlib.import_module("pyPanSim")


def config1_args(n_agents=100_000, n_locations=20_000, weeks=2, seed=0, extra: Sequence[str] = ()) -> List[str]:
    """The CPU-runnable reference workload: dummy population -n/-N, default
    progression/parameters/agentTypes, quarantinePolicy 3, closures off,
    0.1% initially in state E, dt = 10 min."""
    return [
        "-n", str(n_agents), "-N", str(n_locations), "-w", str(weeks), "-t", "10",
        "--quarantinePolicy", "3", "--enableClosures", "0", "--closures", "inputConfigFiles/noRules.json",
        "-c", "configs/configRandom_E0.001.json", "--seed", str(seed), *extra,
    ]


def synthetic_args(n_agents=10_000_000, weeks=4, seed=0, synth_seed=1, extra: Sequence[str] = ()) -> List[str]:
    """Synthetic household population (schools, workplaces, shops, ...),
    all policies on (testing, quarantine policy 3, closure rules)."""
    return [
        "--synthAgents", str(n_agents), "--synthSeed", str(synth_seed), "--synthInfected", "0.001",
        "-w", str(weeks), "-t", "10", "-k", "0.000374", "--quarantinePolicy", "3",
        "--closures", "inputConfigFiles/closureRules.json", "--seed", str(seed), *extra,
    ]
