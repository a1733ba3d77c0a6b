"""The balance identities used to check full-size device runs, pinned on
the reference's own outputs: every cycle of every fixture (all written by
the unmodified reference CPU path) satisfies them, and a planted mistake in
a fixture is caught."""
import copy

import pytest

from conftest import golden_names, load_golden
from quicksilver_amd import checks


@pytest.mark.parametrize("name", golden_names())
def test_reference_fixtures_satisfy_balance_identities(name):
    g = load_golden(name)
    assert checks.balance_violations(g["cycles"]) == []


def test_planted_mistakes_are_caught():
    g = load_golden("small_cube_short")
    for field, delta in [("absorb", 1), ("scatter", -1), ("census", 1), ("split", 1), ("start", 1)]:
        cycles = copy.deepcopy(g["cycles"])
        cycles[1][field] += delta
        assert checks.balance_violations(cycles), field


def test_coral_fixtures_hold_the_reference_verdicts():
    for name in ("coral2_p1", "coral2_p2"):
        g = load_golden(name)
        assert len(g["coral_verdicts"]) == 4
        assert all(v.startswith("PASS::") for v in g["coral_verdicts"]), g["coral_verdicts"]
