"""SetTimeStep (dof18_halfImplicit_gpu.cuh:97 and the 11/24-DOF twins) away from the default 1e-3:
the HIP solver against the oracle at dt = 5e-4 and 2e-3 for all three models, through the C ABI.

Workload: ramp_sine controls (throttle ramp, sinusoidal steering, no braking) computed on the dt grid,
256 HMMWV vehicles, 4 s; 24-DOF vehicles start at U(0, 20) m/s from the records the product's
initializer builds (the oracle starts from the same records through its initializeTireSus).  Bar:
1e-9 normwise relative per vehicle and field over 8 samples, gear exact."""
import os
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import SolverHalfImplicitGPU, TMeasy
from lfdm_amd import workloads as W
from cases18 import normwise_rel
from conftest import DATA

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dt", [5e-4, 2e-3])
@pytest.mark.parametrize("dof", [11, 18, 24])
def test_time_step_vs_oracle(dof, dt):
    import lfdm_oracle as O

    say(f"[timestep] {dof}-DOF dt {dt}")
    d = os.path.join(DATA, f"{dof}dof", "HMMWV")
    flat = lfdm_amd.params_from_json(dof, f"{d}/vehicle.json", f"{d}/tmeasy.json",
                                     f"{d}/suspension.json" if dof == 24 else None)
    n, horizon, samples = 256, 4.0, 8
    nsteps = int(round(horizon / dt))
    every = nsteps // samples
    ctrl = W.ramp_sine(n, nsteps=nsteps, dt=dt, seed=3, max_steer=flat[W.FLAT[dof]["maxSteer"]])
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.SetTimeStep(dt)
    init = None
    if dof == 24:
        recs = W.initial_speed_records(n, flat, dof, seed=4)
        s.initialize_batch(recs)
        init = O.init24(flat, recs)
        np.testing.assert_array_equal(s.get_states(), init)
    else:
        s.Initialize(s.zero_record(), n)
    s.set_controls_steps(ctrl)
    gpu = []
    for _ in range(samples):
        s.advance(every)
        gpu.append(s.get_states())
    gpu = np.stack(gpu)
    s.close()
    _, ref = O.run(dof, flat, ("steps", ctrl), n, nsteps, init=init, dt=dt, every=every)
    err = normwise_rel(gpu, ref)
    say(f"[timestep] {dof}-DOF dt {dt}: max normwise rel {err.max():.3e}, "
        f"bitwise fraction {np.mean(gpu == ref):.4f}")
    assert np.isfinite(gpu).all()
    assert err.max() <= TOL
    np.testing.assert_array_equal(gpu[..., -1], ref[..., -1])
