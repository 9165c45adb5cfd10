"""The oracle's closure rules (RuleClosure, closurePolicies.h:84-480) pinned
for the "afterDays" condition by what its definition fixes
(closurePolicies.h:117-125): a rule with threshold t and openAfter n is
active on the days d (counted from 0 at the first midnight block) with
t <= d < t + n, and while it is active every location of the listed types
is closed, except the essential ones (closurePolicies.h:425-470), and only
those.  Checked one step after each midnight block over
twelve days, for rules with different windows and type lists, from rule
files written here."""
import json

import numpy as np
import pytest

import pansim_amd
from oracle.oracle import OracleSim

SPD = 144


@pytest.mark.parametrize("threshold,open_after,types", [(3, 4, [5, 6]), (0, 2, [4]), (5, 1, [7, 44]),
                                                        (2, 6, [3, 33])])
def test_after_days_rule_closes_exactly_its_types_in_its_window(tmp_path, threshold, open_after, types):
    rules = {"rules": [{"name": "window", "conditionType": "afterDays", "threshold": threshold,
                        "threshold2": threshold, "openAfter": open_after, "closeAfter": -1, "parameter": 0,
                        "locationTypesToClose": types}]}
    path = tmp_path / "rules.json"
    path.write_text(json.dumps(rules))
    args = pansim_amd.synthetic_args(5000, 2, 3, 2, extra=["--closures", str(path), "--enableClosures", "1"])
    o = OracleSim(pansim_amd.with_data_dir(args))
    try:
        lt = o.location_types()
        listed = np.isin(lt, types) & (o.field("locEssential") != 1)
        assert listed.sum() > 0
        for day in range(12):
            o.steps(day * SPD + 1 - o.timestamp)
            closed = o.field("locOpen") == 0
            active = threshold <= day < threshold + open_after
            if active:
                assert np.array_equal(closed, listed), (day, int(closed.sum()), int(listed.sum()))
            else:
                assert not closed.any(), (day, np.unique(lt[closed]))
    finally:
        o.close()
