"""The hand-written 18-DOF step adjoint (csrc/d18_adjoint.h), compiled for the host, against
fourth-order central finite differences of the same step code: gradients of a random linear
functional of (x, y, u, wz) over 300 and 2,000 steps w.r.t. the initial carried state, the controls
and the per-vehicle parameters, normwise relative error <= 1e-6 per group (tests/cpp/adjoint_check.cpp,
built by `make -C low-fidelity-dynamic-models_amd adjcheck`).  Vehicles: HMMWV (torque converter,
4WD), Sedan RWD and audi (no converter).  Over 2,000 steps the braking audi's finite differences
themselves scatter by ~1e-6 (the FD estimate moves by that much between steps 3e-6 and 2e-5), so it
is checked over 300 steps."""
import os
import subprocess

import pytest

from conftest import DATA, ROOT

BIN = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "adjoint_check")


@pytest.mark.parametrize("veh,tire,steps", [("HMMWV/vehicle.json", "HMMWV/tmeasy.json", 300),
                                            ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", 2000),
                                            ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", 300),
                                            ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", 2000),
                                            ("audi/vehicle.json", "audi/tmeasy.json", 300)])
def test_step_adjoint_matches_finite_differences(veh, tire, steps):
    assert os.path.exists(BIN), "run make -C low-fidelity-dynamic-models_amd adjcheck"
    d = os.path.join(DATA, "18dof")
    r = subprocess.run([BIN, f"{d}/{veh}", f"{d}/{tire}", str(steps)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "ALL OK" in r.stdout


LANES = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "adjoint_lanes_check")


@pytest.mark.parametrize("veh", ["HMMWV/vehicle.json", "Sedan/vehicle_RWD.json", "audi/vehicle.json"])
def test_kernel_lane_code_matches_finite_differences(veh):
    """The differentiation-form kernels' lane code (csrc/d18_adjoint_lanes.h: checkpointing,
    segment recompute, buffer indexing, seeding) run on the host over buffers laid out as on the
    device, against fourth-order FD of the same code (tests/cpp/adjoint_lanes_check.cpp)."""
    assert os.path.exists(LANES), "run make -C low-fidelity-dynamic-models_amd adjcheck"
    d = os.path.join(DATA, "18dof")
    tire = f"{d}/{veh.split('/')[0]}/tmeasy.json"
    r = subprocess.run([LANES, f"{d}/{veh}", tire], capture_output=True, text=True, timeout=300)
    print("\n".join(r.stdout.splitlines()[-6:]))
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr
    assert "ALL OK" in r.stdout
