"""Hand-built populations for pinning the movement and planning routines:
a few places, a few agents and agent-type schedules written by the test
itself, in the reference's JSON input formats (agents `people`, locations
`places`, agent types `schedulesUnique` lists: agentTypes.json,
inputRealExample/*.json).  Everything that could move an agent besides its
schedule is switched off unless a test asks for it: infection coefficient 0,
no testing (all testing probabilities 0), no non-COVID hospitalisation, no
closure rules, quarantine policy 0.

Location types are those of the shipped locationTypes.json: public space 1,
home 2, school 3, work 4, hospital 12, doctor 14, classroom 33.
Days: --startDay picks the weekday of day 0 (0 = Monday).
"""
from __future__ import annotations

import json
import os
import tempfile
from typing import Dict, List, Sequence

import numpy as np

PUBLIC, HOME, SCHOOL, WORK, HOSPITAL, DOCTOR, CLASSROOM = 1, 2, 3, 4, 12, 14, 33
STEPS_PER_DAY = 144


def event(loc_type: int, chance: float, start: float, end: float, duration: float) -> Dict:
    """One schedule entry; times in the reference's h.mm notation (8.3 = 8:30)."""
    return {"type": loc_type, "chance": chance, "start": start, "end": end, "duration": duration}


def agent_type(type_id: int, schedules: Sequence[Dict]) -> Dict:
    return {"name": f"type{type_id}", "ID": type_id, "schedulesUnique": list(schedules)}


def schedule(wb: str, day_type: str, events: Sequence[Dict]) -> Dict:
    return {"WB": wb, "dayType": day_type, "schedule": list(events)}


def place(pid: str, loc_type: int, state: str = "ON", essential: int = 0, infectious: float = 1.0) -> Dict:
    return {"ID": pid, "type": loc_type, "essential": essential, "infectious": infectious, "state": state,
            "capacity": 100, "coordinates": [0.0, 0.0], "area": 1, "ageInter": [0, 100]}


def person(type_id: int, locations: Sequence[tuple], state: str = "S", age: int = 40, diagnosed: bool = False,
           pre_cond: str = "0", sex: str = "M") -> Dict:
    return {"age": age, "sex": sex, "preCond": pre_cond, "state": state, "typeID": type_id, "diagnosed": diagnosed,
            "locations": [{"typeID": t, "locID": lid} for t, lid in locations]}


class MiniWorld:
    """Writes the input files into a fresh temporary directory and returns
    the command line of a run over them (args())."""

    def __init__(self, places: List[Dict], people: List[Dict], types: List[Dict]):
        self.dir = tempfile.mkdtemp(prefix="pansim_mini_")
        self.places, self.people, self.types = places, people, types
        self.index = {p["ID"]: i for i, p in enumerate(places)}
        self.cemetery = len(places)
        for name, obj in (("locations.json", {"places": places}), ("agents.json", {"people": people}),
                          ("agentTypes.json", {"types": types})):
            with open(os.path.join(self.dir, name), "w") as fh:
                json.dump(obj, fh)

    def path(self, name):
        return os.path.join(self.dir, name)

    def args(self, weeks=1, start_day=0, quarantine_policy=0, seed=0, extra: Sequence[str] = ()) -> List[str]:
        return ["-a", self.path("agents.json"), "-l", self.path("locations.json"), "-A", self.path("agentTypes.json"),
                "-w", str(weeks), "-t", "10", "-k", "0", "--otherDisease", "0",
                "--testingProbabilities", "0,0,0,0,0,0", "--enableClosures", "0",
                "--closures", "inputConfigFiles/noRules.json", "--quarantinePolicy", str(quarantine_policy),
                "--startDay", str(start_day), "--seed", str(seed), *extra]

    def loc(self, pid: str) -> int:
        return self.index[pid]


def trajectory(sim, n_steps: int, fields=("location",)):
    """Steps `sim` one step at a time; returns {field: array[n_steps, n_agents]}
    of the values after each step."""
    out = {f: [] for f in fields}
    for _ in range(n_steps):
        sim.steps(1)
        for f in fields:
            out[f].append(sim.field(f).copy())
    return {f: np.array(v) for f, v in out.items()}


def runs(spec: Sequence[tuple]) -> np.ndarray:
    """Expands [(location, first_step, last_step), ...] (inclusive) into a
    per-step array."""
    n = max(e for _, _, e in spec) + 1
    a = np.full(n, -1, dtype=np.int64)
    for loc, b, e in spec:
        a[b:e + 1] = loc
    assert (a >= 0).all(), "gap in the expected trajectory"
    return a
