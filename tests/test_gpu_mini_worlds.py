"""GPU parity on the hand-built populations whose oracle trajectories are
worked out by hand in tests/test_oracle_movement_pins.py and
tests/test_oracle_schedule_actions.py: the HIP path must reproduce them step
by step (every field after every step of the checked days)."""
import pytest

import pansim_amd
from tests import test_oracle_movement_pins as pins
from tests import test_oracle_schedule_actions as actions
from tests.parity_util import run_parity

pytestmark = pytest.mark.gpu


def _every_step(n):
    return list(range(1, n + 1))


def test_mini_schedule_day():
    t = pins.agent_type(1, [pins.schedule("W", "ALL", [
        pins.event(pins.WORK, 1, 8, 8.1, 2), pins.event(pins.SHOP7, 1, 10.4, 10.5, 0.3),
        pins.event(pins.SHOP8, 1, 11.3, 11.4, 1), pins.event(pins.SHOP7, 1, 14, 14.1, 0.3)])])
    w = pins.MiniWorld(pins._world_places(), [pins.person(1, pins.ALL_LOCS)], [t])
    run_parity(w.args(), _every_step(146))


@pytest.mark.parametrize("qp", [0, 1, 2, 3])
def test_mini_quarantine(qp):
    run_parity(pins._quarantine_world().args(quarantine_policy=qp), _every_step(150))


@pytest.mark.parametrize("qp,g2", [(3, "I1"), (4, "I1"), (4, "S")])
def test_mini_school_planning(qp, g2):
    run_parity(pins._school_world(g2_state=g2).args(quarantine_policy=qp), [48, 100, 144, 145, 195, 288, 300])


def test_mini_holiday_and_reduce_movement():
    t1 = pins.agent_type(1, [pins.schedule("W", "ALL", [
        pins.event(actions.SHOP7, .5, 8, 8.1, 1), pins.event(actions.SHOP8, .5, 8, 8.1, 1),
        pins.event(pins.WORK, 1, 12, 12.1, 1)])])
    t2 = pins.agent_type(2, [pins.schedule("W", "WEEKDAYS", [pins.event(actions.SHOP8, 1, 8, 17, 1)])])
    w = pins.MiniWorld(actions._places(), [pins.person(1 + (i % 2), actions.LOCS) for i in range(400)], [t1, t2])
    run_parity(w.args(extra=actions._rules(w, "ReduceMovement", 1, 2, "0.5")), [49, 144, 193, 337, 481, 600])
    w2 = pins.MiniWorld(actions._places(), [pins.person(1 + (i % 2), actions.LOCS) for i in range(50)], [t1, t2])
    run_parity(w2.args(extra=actions._rules(w2, "HolidayMode", 1, 1)), [49, 144, 193, 250, 337])
