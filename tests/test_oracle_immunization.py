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


PROT = [0.7, 0.3, 0.6, 0.4, 0.8, 0.5, 0.9, 0.2, 0.85, 0.35]
WANE = [0.01, 0.005, 0.02, 0.02, 0.004, 0.003, 0.006, 0.001, 0.002, 0.007]
GROUP = [0, 0, 0, 1, 1, 1, 1]  # --variantSimilarity default


def _vs(pair, weeks, same_variant):
    """-W * weeks + (0.95 for the variant of the earlier infection, else P):
    float products, the sum in double when the 0.95 literal is involved."""
    w, p = np.float32(WANE[pair]), np.float32(PROT[pair])
    lin = np.float32(-1.0) * w * np.float32(weeks)
    if same_variant is None:
        return np.float32(lin + p)
    return np.float32(float(lin) + (0.95 if same_variant else float(p)))


def test_protection_after_infection_and_boosters_follows_the_rule():
    """The other cases of the same pass, with every protection parameter set
    to a distinct value and every waning rate non-zero:
      infected before, not vaccinated (or one dose less than 4 weeks old):
        per variant v, the pair (protectionInfection[k], waning[k]) with
        k = 0 when an earlier infection was of a variant of v's similarity
        group, else 1; the variant itself of an earlier infection: 0.95;
        weeks since the infection;
      vaccinated and infected before: pairs 6/7 (one dose) or 8/9 (more),
        weeks = the fewer of the two;
      vaccinated, never infected, more than one dose: pairs 4/5 (group 0 =
        similar to the vaccine), weeks since the last dose;
    susceptibility = min(1, 1 - (-W * weeks + P)).  S agents are recounted
    over 20 weeks (infections with several variants, recoveries back to S,
    first doses from day 2 and boosters)."""
    args = pansim_amd.synthetic_args(4000, 20, 6, 3, extra=[
        "-k", "0.003", "--synthInfected", "0.05", "--immunizationStart", "2", "--immunizationsPerWeek", "700",
        "--boosterStart", "40",
        "--protectionInfection", ",".join(map(str, PROT)), "--protectionInfectionWaning", ",".join(map(str, WANE)),
        "--closures", "inputConfigFiles/noRules.json"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    n_var = len(o.variant_multipliers())
    cases = {"inf": 0, "inf+1": 0, "inf+more": 0, "more": 0}
    try:
        for day in range(35, 140, 5):
            # the pass reads the histories as they are before the midnight
            # block's own doses (it runs before the rollout, immunization.h:481-900)
            ts = day * SPD
            o.steps(ts - o.timestamp)
            cnt = o.field("infectedCount").astype(np.int64)
            imm = o.field("immunizationCount").astype(np.int64)
            its = o.field("infectedTimestamp").astype(np.int64)
            mts = o.field("immunizationTimestamp").astype(np.int64)
            prior = o.field("statsVariant").astype(np.int64)
            o.steps(1)
            state = o.field("state")
            sus = [o.field(f"susceptible{v}").view(np.float32) for v in range(n_var)]
            for i in np.nonzero((state == 0) & ((cnt > 0) | (imm > 1)))[0]:
                w_inf = (ts - its[i]) // WEEK
                w_imm = (ts - mts[i]) // WEEK
                for v in range(n_var):
                    similar = any((prior[i] >> j) & 1 and GROUP[j] == GROUP[v] for j in range(n_var))
                    same = bool((prior[i] >> v) & 1)
                    if cnt[i] > 0 and (imm[i] == 0 or (imm[i] == 1 and w_imm < 4)):
                        vs, case = _vs(0 if similar else 1, w_inf, same), "inf"
                    elif cnt[i] > 0:
                        base = 6 if imm[i] == 1 else 8
                        vs = _vs(base + (0 if similar else 1), min(w_inf, w_imm), same)
                        case = "inf+1" if imm[i] == 1 else "inf+more"
                    else:
                        vs, case = _vs(4 + (0 if GROUP[v] == 0 else 1), w_imm, None), "more"
                    expect = min(np.float32(1.0), np.float32(1.0) - vs)
                    assert abs(float(sus[v][i]) - float(expect)) <= 1e-6, (day, i, v, case, sus[v][i], expect)
                    cases[case] += 1
    finally:
        o.close()
    assert all(c > 20 for c in cases.values()), cases
