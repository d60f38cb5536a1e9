"""Failure detection helpers on the host (batched_simulate.py:158-205)."""

from __future__ import annotations

import numpy as np
import pytest

from prolix_amd.checks import SimulationInstability, check_position_sanity


def test_position_sanity_names_every_failed_system():
    rng = np.random.default_rng(0)
    sizes = [5, 7, 4]
    x = rng.uniform(-10, 10, size=(16, 3)).astype(np.float32)
    check_position_sanity(x, sizes, 50.0, "equilibration")  # all sane
    x[6, 1] = 120.0   # system 1 exceeds
    x[13, 0] = np.nan  # system 2 is non-finite
    with pytest.raises(SimulationInstability) as ei:
        check_position_sanity(x, sizes, 50.0, "production", system_names=["a", "b", "c"])
    assert ei.value.systems == [1, 2]
    assert "production" in str(ei.value) and "b=120.0" in str(ei.value)


def test_position_sanity_bound_is_inclusive_of_the_limit():
    x = np.zeros((3, 3), np.float32)
    x[1, 2] = 50.0
    check_position_sanity(x, [3], 50.0, "minimization")
    x[1, 2] = np.nextafter(np.float32(50.0), np.float32(60.0))
    with pytest.raises(SimulationInstability):
        check_position_sanity(x, [3], 50.0, "minimization")
