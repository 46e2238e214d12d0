"""Batching-form host logic that needs no GPU: argument checks of the horizon, layout and fleet entry
points, and their errors before any device use."""
import os

import numpy as np
import pytest

from lfdm_amd import LAYOUT_BY_HORIZON, LAYOUT_GROUPED, Fleet, LfdmError, SolverHalfImplicitGPU
from conftest import DATA


def solver(n=10):
    d = os.path.join(DATA, "18dof")
    s = SolverHalfImplicitGPU(18, n)
    s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", n, os.path.join(DATA, "input", "acc3.txt"))
    return s


def test_horizon_and_layout_arguments():
    s = solver()
    assert s.get_layout() == LAYOUT_GROUPED
    with pytest.raises(LfdmError) as e:
        s.set_horizons(np.array([5] * 9 + [-1]))
    assert e.value.name == "LFDM_E_ARG"
    with pytest.raises(ValueError):
        s.set_horizons(np.arange(3))
    with pytest.raises(LfdmError) as e:
        s.set_layout(7)
    assert e.value.name == "LFDM_E_ARG"
    # before the device image exists these only record the request
    s.set_horizons(np.arange(10) * 100)
    s.set_layout(LAYOUT_BY_HORIZON)
    assert s.get_layout() == LAYOUT_BY_HORIZON
    s.set_horizons(None)
    s.close()


def test_fleet_arguments():
    f = Fleet()
    with pytest.raises(LfdmError) as e:  # nothing to run
        f.run(10)
    assert e.value.name == "LFDM_E_STATE"
    with pytest.raises(LfdmError) as e:
        f.last_ms()
    assert e.value.name == "LFDM_E_STATE"
    s = solver()
    assert f.add(s) == 0
    with pytest.raises(LfdmError) as e:  # a handle is one bucket
        f.add(s)
    assert e.value.name == "LFDM_E_ARG"
    for bad in ((1, 256, 0), (0, 96, 0), (0, 512, 0), (0, 256, -1), (0, 256, 0, 4)):
        with pytest.raises(LfdmError) as e:
            f.set_config(*bad)
        assert e.value.name == "LFDM_E_ARG"
    f.set_config(0, 128, 500, 1)
    assert f.get_config(0) == (128, 500, 1, LAYOUT_GROUPED)
    with pytest.raises(LfdmError) as e:
        f.autotune(100, trial_steps=0)
    assert e.value.name == "LFDM_E_ARG"
    # a bucket whose vehicles are not all initialized fails before any device use
    with pytest.raises(LfdmError) as e:
        f.run(10)
    assert e.value.name == "LFDM_E_STATE"
    f.close()
    s.close()
