"""The scheduler's switching rule (dyng_amd.scheduler.pick, the same rule as pick_schedule in
dyng_amd/csrc/scheduler.hip): the plain schedule (0) runs unless another is predicted at least
15 % faster and at least 20 us sooner. Host logic only; no device."""

from __future__ import annotations

import math

import pytest

from dyng_amd import scheduler


def ms(*values):
    return [math.log2(v) for v in values]


@pytest.mark.parametrize("predicted, want", [
    (ms(1.0, 0.80, 0.90, 0.95, 2.0), 1),    # 20 % and 200 us faster: switch
    (ms(1.0, 0.90, 0.95, 0.88, 2.0), 0),    # the best takes 88 % of the time, 13.6 % faster: stay
    (ms(1.0, 0.90, 0.95, 0.86, 2.0), 3),    # 86 % of the time is 16.3 % faster: switch
    (ms(0.05, 0.04, 0.06, 0.07, 0.08), 0),  # 20 % faster but only 10 us sooner: stay
    (ms(0.20, 0.30, 0.10, 0.40, 0.50), 2),  # half the time and 100 us sooner: switch
    (ms(0.5, 0.6, 0.7, 0.8, 0.9), 0),       # the plain schedule is the fastest
    (ms(1.0, 1.0, 1.0, 1.0, 1.0), 0),       # a tie keeps the plain schedule
])
def test_pick(predicted, want):
    assert scheduler.pick(predicted) == want


def test_the_margin_is_fifteen_percent():
    assert scheduler.SWITCH_MARGIN == pytest.approx(math.log2(1.15), abs=1e-4)
    assert scheduler.MIN_GAIN_MS == pytest.approx(0.02)
    # just past the margin and the floor on both sides
    assert scheduler.pick(ms(1.0, 1.0 / 1.151, 2, 2, 2)) == 1
    assert scheduler.pick(ms(1.0, 1.0 / 1.149, 2, 2, 2)) == 0
    # past the margin (17 % and 19 % faster), on both sides of the 20 us floor
    assert scheduler.pick(ms(0.13, 0.109, 1, 1, 1)) == 1  # 21 us sooner
    assert scheduler.pick(ms(0.13, 0.111, 1, 1, 1)) == 0  # 19 us sooner


def test_the_rule_matches_the_library_source():
    """The constants of the device-side rule are the same as the host-side ones."""
    import pathlib
    import re

    src = (pathlib.Path(__file__).resolve().parents[1] / "dyng_amd" / "csrc" / "scheduler.hip").read_text()
    margin = float(re.search(r"kSwitchMargin = ([0-9.]+)f", src).group(1))
    floor = float(re.search(r"kMinGainMs = ([0-9.]+)f", src).group(1))
    assert margin == pytest.approx(scheduler.SWITCH_MARGIN) and floor == pytest.approx(scheduler.MIN_GAIN_MS)
