"""The reference's GPU known-answer tests and SolveStep(t, functor) demo, compiled as C++ against the
drop-in classes of include/lfdm_amd.hpp (tests/cpp/dropin_test.cpp, built by
`make -C low-fidelity-dynamic-models_amd dropin`)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import DATA, ROOT

BIN = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "dropin_test")


def test_dropin_program_is_built():
    assert os.path.exists(BIN), "run make -C low-fidelity-dynamic-models_amd dropin"


def functor_controls(n, tend=10.0, ksim=0.1, dt=1e-3):
    """Per-step controls (steps, 3, n) of the program's per-vehicle functor, at the times its SolveStep
    chunks evaluate it: chunk k starts at t_k (t_0 = 0, t_k+1 = t_k + ksim, as SolveStep returns), its steps
    at t_k, t_k + dt, ... (accumulated), while t < t_k + ksim - dt / 10 (dof18_halfImplicit_gpu.cu:892)."""
    times = []
    t0 = 0.0
    while t0 < tend - 1e-4:
        t, end = t0, (t0 + ksim) - dt / 10.0
        while t < end:
            times.append(t)
            t = t + dt
        t0 = t0 + ksim
    v = np.arange(n)
    c = np.zeros((len(times), 3, n))
    for k, t in enumerate(times):
        c[k, 0] = 0.0 if t < 1.0 else 0.0625 * (v % 9) - 0.25
        c[k, 1] = 0.0 if t < 0.5 else 0.25 + 0.0625 * (v % 11)
        c[k, 2] = np.where((v % 13 == 0) & (t > 6.0), 0.375, 0.0)
    return c


@pytest.mark.gpu
def test_dropin_program_passes(tmp_path):
    """The program's checks, then its per-vehicle functor fleet (controls from the thread index, evaluated
    per lane and step) against the oracle fed the same per-step controls, final records under the fleet bar
    of fleetcheck.py (1e-9 of the trajectory's scale for ulp-stable vehicles; the braking vehicles lock
    wheels, where one ulp shows)."""
    import fleetcheck as F
    import lfdm_amd
    import lfdm_oracle as O
    from lfdm_amd import workloads as W

    out = tmp_path / "functor_records.bin"
    r = subprocess.run([BIN, DATA, str(out)], capture_output=True, text=True, timeout=600)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
    n = 200
    gpu = np.fromfile(out, dtype=np.float64).reshape(n, 75)
    flat = lfdm_amd.params_from_json(18, f"{DATA}/18dof/HMMWV/vehicle.json", f"{DATA}/18dof/HMMWV/tmeasy.json")
    ctrl = functor_controls(n)
    R = F.final_reference(O, 18, flat, ("steps", ctrl), n, ctrl.shape[0], W.default_overrides(flat, n))
    v = F.final_verdict(gpu, R)
    print(F.summary_line("per-vehicle functor fleet vs oracle", v), flush=True)
    assert v["ok"], v["failing"]
    assert len(set(np.round(R.traj[:, 60], 6))) > 50  # the controls really differ per vehicle
