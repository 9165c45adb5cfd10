"""The oracle's immunization pass (Immunization::update, immunization.h:
547-695), case "vaccinated, never infected", pinned against the rule's
parameters as given on the command line: for an agent in state S with one
dose, w whole weeks after it, the susceptibility to variant 0 (vaccines are
of variant type 0, so the "similar" 2-dose entry applies) is

    1                          for w = 0 (no protection in the first week),
    1 - (P - W * w) / 2        for 1 <= w <= 3 (half protection),
    1 - (P - W * w)            from w = 4 (waning protection),

with P = protectionInfection[2] and W = protectionInfectionWaning[2],
chosen here unlike their defaults so that a hard-wired value fails; agents
never vaccinated nor infected stay fully susceptible.  The agents' dose
counts and dates are read from their fields at several midnights."""
import numpy as np

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144
WEEK = 7 * SPD
P = np.float32(0.6)
W = np.float32(0.02)


def test_first_dose_protection_follows_the_rule():
    prot = "0.85,0.5,0.6,0.4,0.85,0.63,0.9,0.65,0.92,0.7"
    waning = "0.003,0.006,0.02,0.02,0.005,0.02,0.002,0.004,0.002,0.004"
    args = pansim_amd.config1_args(1500, 300, 12, 4, extra=[
        "-k", "0.0", "--immunizationStart", "2", "--immunizationsPerWeek", "700",
        "--protectionInfection", prot, "--protectionInfectionWaning", waning])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        seen = set()
        for day in (3, 9, 16, 23, 30, 37, 44, 51, 58, 65, 72):
            o.steps(day * SPD + 1 - o.timestamp)
            ts = day * SPD  # the midnight block that just ran
            state = o.field("state")
            sus = o.field("susceptible0").view(np.float32)
            count = o.field("immunizationCount")
            when = o.field("immunizationTimestamp")
            never_inf = o.field("infectedCount") == 0
            plain = (state == 0) & never_inf & (count == 0)
            assert np.all(sus[plain] == 1.0)
            one = (state == 0) & never_inf & (count == 1)
            for i in np.nonzero(one)[0]:
                w = (ts - int(when[i])) // WEEK
                vs = np.float32(-1.0) * W * np.float32(w) + P
                if w == 0:
                    expect = np.float32(1.0)
                elif w < 4:
                    expect = np.float32(1.0) - vs / np.float32(2.0)
                else:
                    expect = np.float32(1.0) - vs
                assert abs(float(sus[i]) - float(expect)) <= 1e-6, (day, i, w, sus[i], expect)
                seen.add(min(int(w), 4))
        assert seen == {0, 1, 2, 3, 4}  # every branch of the rule was reached
    finally:
        o.close()
