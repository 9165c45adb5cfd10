"""The reference's example population, inputRealExample/agents.tar and
locations.tar, committed unchanged under tests/golden/real_example/: 22,364
agents (16,365 once the tourists are skipped) with 19 possible locations each
on average, 8,165 places.  The file has every agent in state S; the runs here
start a fraction of them, drawn by a fixed generator, in state E."""
from __future__ import annotations

import json
import os
import random
import tarfile
from typing import List

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "real_example")


def _member(name: str) -> bytes:
    with tarfile.open(os.path.join(HERE, name + ".tar")) as t:
        f = t.extractfile(name + ".json")
        assert f is not None, name
        return f.read()


def write_inputs(directory, infected: float = 0.01, seed: int = 1) -> List[str]:
    """Writes agents.json (with the exposed agents) and locations.json into
    `directory`; returns the input-file part of a command line."""
    agents = json.loads(_member("agents"))
    rng = random.Random(seed)
    for p in agents["people"]:
        if rng.random() < infected:
            p["state"] = "E"
    a = os.path.join(str(directory), "agents.json")
    l = os.path.join(str(directory), "locations.json")
    with open(a, "w") as fh:
        json.dump(agents, fh)
    with open(l, "wb") as fh:
        fh.write(_member("locations"))
    return ["-a", a, "-l", l]


def args(directory, weeks: int = 2, seed: int = 1, k: str = "0.000374") -> List[str]:
    """The example population with the shipped schedules, progression and
    closure rules, quarantine policy 3, testing on."""
    return [*write_inputs(directory), "-w", str(weeks), "-t", "10", "-k", k, "--quarantinePolicy", "3",
            "--closures", "inputConfigFiles/closureRules.json", "--seed", str(seed)]
