"""11- and 24-DOF HIP solver parity on the GPU, through the C ABI.

Same bar as the 18-DOF tests: trajectories within 1e-9 normwise-relative of the reference (max over
time of |a - b| / max over time of |b|, per vehicle and field) on every golden case, integer gear
exact (fleet-scale parity: test_gpu_fleets.py).  The 24-DOF cases start from the records the
product's Initialize builds (initializeTireSus), so the initializer is exercised too.
"""
import os
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import SolverHalfImplicitGPU, TMeasy, TMeasyNr
from lfdm_amd import workloads as W
from cases1124 import CASES, KAT_X, X_FROM_END, inputs, paths
from cases18 import final_rel, normwise_rel
from conftest import DATA

pytestmark = pytest.mark.gpu

TOL = 1e-9
IDS = [(dof, name) for dof in (11, 24) for name in sorted(CASES[dof])]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden(request, dof):
    return request.getfixturevalue(f"golden{dof}")


def make_solver(dof, name, g):
    veh, tire, nr, kind, drv, sus = paths(dof, name)
    flat, ctrl, n, ovr, init = inputs(dof, name, g)
    tt = TMeasyNr if nr else TMeasy
    s = SolverHalfImplicitGPU(dof, n)
    rows = None if drv is None else lfdm_amd.load_driver_file(drv)
    s.construct_flat(flat, n, rows, tt)
    if ovr is not None:
        s.set_overrides(ovr)
    rec = s.zero_record()
    if dof == 24:
        rec[-24 + 3] = g[f"{name}/spec"][3]  # initial forward speed of the case
    s.Initialize(rec, n)
    if kind == "knots":
        _, ns, ds, np_, dp, knots = ctrl
        s.set_controls_knots(ns, ds, np_, dp, knots)
    elif kind == "step_api":
        s.set_controls_steps(ctrl[1])
    return s


def run_sampled(s, nsteps, every):
    out = []
    for _ in range(nsteps // every):
        s.advance(every)
        out.append(s.get_states())
    return np.stack(out)


@pytest.mark.parametrize("dof,name", IDS)
def test_golden_trajectories(dof, name, request):
    g = golden(request, dof)
    say(f"[gpu{dof}] golden case {name}")
    nsteps, every = (int(v) for v in g[f"{name}/spec"][:2])
    s = make_solver(dof, name, g)
    np.testing.assert_array_equal(s.get_states(), g[f"{name}/init"])
    traj = run_sampled(s, nsteps, every)
    ref = g[f"{name}/traj"]
    err = normwise_rel(traj, ref)
    say(f"[gpu{dof}] {name}: max normwise rel {err.max():.3e}, bitwise {np.array_equal(traj, ref)}")
    assert err.max() <= TOL
    np.testing.assert_array_equal(traj[..., -1], ref[..., -1])
    fin = s.get_states()
    assert final_rel(fin, ref, g[f"{name}/final"]).max() <= TOL
    if name in KAT_X[dof]:
        assert abs(fin[0, X_FROM_END[dof]] - KAT_X[dof][name]) < 1.0
    s.close()


@pytest.mark.parametrize("dof", [11, 24])
def test_reference_gpu_fleet_test(dof, request):
    """tests/11dof_gpu.cu:43-82 and tests/24dof_gpu.cu:62-110: 500 vehicles on acc3 and 500 on
    double_lane4, 22 s, two Construct calls, threads per block 32; vehicle 499 and 999 end near the
    reference's values, and match the golden 22 s trajectories."""
    g = golden(request, dof)
    say(f"[gpu{dof}] two-group fleet, 1000 vehicles, 22 s")
    d = os.path.join(DATA, f"{dof}dof", "HMMWV")
    s = SolverHalfImplicitGPU(dof, 1000)
    extra = [f"{d}/suspension.json"] if dof == 24 else []
    for drv in ("acc3", "double_lane4"):
        s.Construct(f"{d}/vehicle.json", f"{d}/tmeasy.json", *extra, 500, os.path.join(DATA, "input", f"{drv}.txt"))
    s.SetThreadsPerBlock(32)
    s.SetTimeStep(1e-3)
    s.Initialize(s.zero_record(), 1000)
    s.SetEndTime(22.0)
    s.Solve()
    xi = X_FROM_END[dof]
    assert abs(s.GetSimState(499)[xi] - KAT_X[dof]["hmmwv_acc3_22s"]) < 1
    assert abs(s.GetSimState(999)[xi] - KAT_X[dof]["hmmwv_dl4_22s"]) < 1
    allst = s.get_states()
    for lo, hi, case in ((0, 500, "hmmwv_acc3_22s"), (500, 1000, "hmmwv_dl4_22s")):
        assert final_rel(allst[lo:hi], g[f"{case}/traj"][:, 0], g[f"{case}/final"][0]).max() <= TOL
        assert np.all(allst[lo:hi] == allst[lo])
    s.close()


@pytest.mark.parametrize("dof", [11, 24])
def test_csv_output_format(dof, tmp_path, request):
    """SetOutput writes <file>_<vehicle>.csv with the model's header, a zero first row and one row per
    output sample (dof11_halfImplicit_gpu.cu:555-610, dof24_halfImplicit_gpu.cu:600-670)."""
    g = golden(request, dof)
    say(f"[gpu{dof}] CSV output")
    d = os.path.join(DATA, f"{dof}dof", "HMMWV")
    extra = [f"{d}/suspension.json"] if dof == 24 else []
    s = SolverHalfImplicitGPU(dof, 100)
    s.Construct(f"{d}/vehicle.json", f"{d}/tmeasy.json", *extra, 100, os.path.join(DATA, "input", "acc3.txt"))
    s.Initialize(s.zero_record(), 100)
    s.SetEndTime(10.0)
    base = str(tmp_path / "run")
    s.SetOutput(base, 10, False, 5)
    s.Solve()
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("run_"))
    assert len(files) == 5
    text = open(os.path.join(tmp_path, files[0])).read().splitlines()
    cols = lfdm_amd.OUTPUT_COLUMNS[dof]
    assert text[0] == ",".join(cols) + ","
    assert text[1] == ",".join(["0"] * len(cols)) + ","
    assert len(text) == 2 + 100
    last = [float(v) for v in text[-1].rstrip(",").split(",")]
    assert abs(last[0] - 10.0) < 1e-6
    ref = g["hmmwv_acc3/final"][0]
    assert abs(last[1] - ref[X_FROM_END[dof]]) <= 1e-4 * abs(ref[X_FROM_END[dof]])
    s.close()
