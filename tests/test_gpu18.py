"""18-DOF HIP solver parity on the GPU, through the C ABI.

Bar: trajectories within 1e-9 normwise-relative of the reference (max over time of |a - b| divided
by max over time of |b|, per vehicle and state field) on every golden case, and within 1e-9 of the
oracle on the 1,024-vehicle ramp_sine workload.  Integer-like and exact properties (gear,
batch independence, chunking invariance) are checked bit for bit.  Fleet-scale parity of all models:
test_gpu_fleets.py.
"""
import os
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import SolverHalfImplicitGPU, TMeasy, TMeasyNr
from lfdm_amd import workloads as W
from cases18 import CASES, KAT_X, final_rel, inputs, normwise_rel, paths
from conftest import DATA

pytestmark = pytest.mark.gpu

TOL = 1e-9
DCRANK = -2  # dOmega_crank: never computed without a torque converter


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_solver(name, g, n=None):
    veh, tire, nr, kind, drv = paths(name)
    flat, ctrl, n_case, ovr = inputs(name, g)
    n = n_case if n is None else n
    tt = TMeasyNr if nr else TMeasy
    s = SolverHalfImplicitGPU(18, n)
    rows = None if drv is None else lfdm_amd.load_driver_file(drv)
    s.construct_flat(flat, n, rows, tt)
    s.Initialize(s.zero_record(), n)
    if kind == "knots":
        _, ns, ds, np_, dp, knots = ctrl
        s.set_controls_knots(ns, ds, np_, dp, knots)
        s.set_overrides(ovr)
    elif kind in ("step_api", "ramp_sine"):
        s.set_controls_steps(ctrl[1])
    return s


def run_sampled(s, nsteps, every):
    out = []
    for _ in range(nsteps // every):
        s.advance(every)
        out.append(s.get_states())
    return np.stack(out)


def masked(name, g, a):
    a = a.copy()
    if g[f"{name}/params"][21] == 0:
        a[..., DCRANK] = 0.0
    return a


@pytest.mark.parametrize("name", sorted(CASES))
def test_golden_trajectories(name, golden18):
    say(f"[gpu18] golden case {name}")
    nsteps, every, nr = (int(v) for v in golden18[f"{name}/spec"])
    s = make_solver(name, golden18)
    traj = masked(name, golden18, run_sampled(s, nsteps, every))
    ref = masked(name, golden18, golden18[f"{name}/traj"])
    err = normwise_rel(traj, ref)
    say(f"[gpu18] {name}: max normwise rel {err.max():.3e}, bitwise {np.array_equal(traj, ref)}")
    assert err.max() <= TOL
    # the gear index is an integer: it must agree exactly
    np.testing.assert_array_equal(traj[..., -1], ref[..., -1])
    fin = masked(name, golden18, s.get_states())
    assert final_rel(fin, ref, masked(name, golden18, golden18[f"{name}/final"])).max() <= TOL
    if name in KAT_X:
        assert abs(fin[0, -15] - KAT_X[name]) < 1.0
    s.close()


def test_reference_gpu_fleet_test(golden18):
    """wheeled_vehicle_models/tests/18dof_gpu.cu:58-99: 500 vehicles on acc3 + 500 on double_lane4,
    22 s, two Construct calls; vehicle 499 -> 241.112, vehicle 999 -> 106.548 (+-1)."""
    say("[gpu18] two-group fleet, 1000 vehicles, 22 s")
    d = os.path.join(DATA, "18dof")
    s = SolverHalfImplicitGPU(18, 1000)
    s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 500, os.path.join(DATA, "input", "acc3.txt"))
    s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 500,
                os.path.join(DATA, "input", "double_lane4.txt"))
    s.SetThreadsPerBlock(32)
    s.SetTimeStep(1e-3)
    s.Initialize(s.zero_record(), 1000)
    s.SetEndTime(22.0)
    s.Solve()
    x499, x999 = s.GetSimState(499)[60], s.GetSimState(999)[60]
    assert abs(x499 - 241.112) < 1 and abs(x999 - 106.548) < 1
    allst = s.get_states()
    for lo, hi, case in ((0, 500, "hmmwv_acc3_22s"), (500, 1000, "hmmwv_dl4_22s")):
        ref = golden18[f"{case}/final"][0]
        assert final_rel(allst[lo:hi], golden18[f"{case}/traj"][:, 0], ref).max() <= TOL
        assert np.all(allst[lo:hi] == allst[lo])  # identical vehicles stay bit-identical
    s.close()


def test_heterogeneous_construct_groups(golden18):
    """Three Construct groups of different vehicles in one handle (HMMWV on acc3, Sedan RWD on acc3,
    audi on double_lane4; 70 / 50 / 90 vehicles, so every group ends in a padded wavefront), solved
    together for 10 s: each group ends where the reference's own run of that vehicle ends (1e-9,
    gear exact; dOmega_crank, which the reference leaves uninitialized without a torque converter,
    masked as in the golden cases), identical vehicles bit-identical."""
    say("[gpu18] heterogeneous Construct groups")
    d = os.path.join(DATA, "18dof")
    inp = os.path.join(DATA, "input")
    groups = (("HMMWV/vehicle.json", "HMMWV/tmeasy.json", "acc3.txt", 70, "hmmwv_acc3"),
              ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", "acc3.txt", 50, "sedan_rwd_acc3"),
              ("audi/vehicle.json", "audi/tmeasy.json", "double_lane4.txt", 90, "audi_dl4"))
    total = sum(g[3] for g in groups)
    s = SolverHalfImplicitGPU(18, total)
    for veh, tire, drv, n, _ in groups:
        s.Construct(f"{d}/{veh}", f"{d}/{tire}", n, os.path.join(inp, drv))
    s.Initialize(s.zero_record(), total)
    s.SetEndTime(10.0)
    s.Solve()
    allst = s.get_states()
    s.close()
    lo = 0
    for _, _, _, n, case in groups:
        got = masked(case, golden18, allst[lo:lo + n])
        ref = masked(case, golden18, golden18[f"{case}/final"][0])
        err = final_rel(got, masked(case, golden18, golden18[f"{case}/traj"][:, 0]), ref).max()
        say(f"[gpu18] {case} group: max rel {err:.3e}")
        assert err <= TOL
        np.testing.assert_array_equal(got[:, -1], np.full(n, ref[-1]))
        assert np.all(got == got[0])
        lo += n


def test_sampled_output_matches_full_output(tmp_path):
    """SetOutput(file, freq, store_all=false, no_outs) writes only the sampled vehicles, whose columns
    the device gathers before the copy to the host: their CSV files are identical, byte for byte, to
    the same vehicles' files of a store_all run.  Two Construct groups (so sampled vehicles differ),
    2 s launches, host dumps every 4 s (mid-run and final flushes)."""
    say("[gpu18] sampled vs full CSV output")
    d = os.path.join(DATA, "18dof")
    names = {}
    for tag, store_all, no_outs in (("full", True, 0), ("some", False, 20)):
        s = SolverHalfImplicitGPU(18, 100)
        s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 60, os.path.join(DATA, "input", "acc3.txt"))
        s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 40,
                    os.path.join(DATA, "input", "double_lane4.txt"))
        s.Initialize(s.zero_record(), 100)
        s.SetEndTime(10.0)
        s.SetKernelSimTime(2.0)
        s.SetHostDumpTime(4.0)
        s.SetOutput(str(tmp_path / tag), 20, store_all, no_outs)
        s.Solve()
        s.close()
        names[tag] = sorted(f for f in os.listdir(tmp_path) if f.startswith(tag + "_"))
    assert len(names["full"]) == 100 and len(names["some"]) == 20
    sampled = [int(f[len("some_"):-len(".csv")]) for f in names["some"]]
    assert min(sampled) < 60 <= max(sampled)  # both groups are sampled
    for v in sampled:
        a = open(tmp_path / f"some_{v}.csv").read()
        b = open(tmp_path / f"full_{v}.csv").read()
        assert a == b, v
        assert len(a.splitlines()) == 2 + 200


def test_ramp_sine_1024_vs_oracle(golden18):
    """The 1,024-vehicle ramp_sine workload (10 s) against the oracle at 1e-9."""
    import lfdm_oracle as O

    say("[gpu18] ramp_sine 1024 vehicles vs oracle")
    flat = golden18["hmmwv_acc3/params"]
    n, nsteps = 1024, 10000
    ctrl = W.ramp_sine(n, nsteps=nsteps, seed=1, max_steer=flat[W.FLAT18["maxSteer"]])
    s = SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_controls_steps(ctrl)
    every = 1000
    traj = run_sampled(s, nsteps, every)
    fin_o, traj_o = O.run18(flat, ("steps", ctrl), n, nsteps, every=every)
    err = normwise_rel(traj, traj_o)
    say(f"[gpu18] ramp_sine: max normwise rel {err.max():.3e}; bitwise fraction "
        f"{np.mean(traj == traj_o):.4f}")
    assert err.max() <= TOL
    np.testing.assert_array_equal(traj[..., -1], traj_o[..., -1])
    s.close()


def test_batch_independence(golden18):
    """A vehicle's trajectory is bit-identical whether it runs in a 4,096-vehicle knot fleet or in a
    fleet of three (fleet-scale parity against the oracle: test_gpu_fleets.py)."""
    say("[gpu18] batch independence")
    flat = golden18["hmmwv_acc3/params"]
    n, nsteps = 4096, 4000
    fl = W.knot_fleet(n, flat, seed=5)
    s = SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_overrides(fl["ovr"])
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    s.advance(nsteps)
    pick = [7, 1000, 4095]
    s3 = SolverHalfImplicitGPU(18, 3)
    s3.construct_flat(flat, 3, None, TMeasy)
    s3.Initialize(s3.zero_record(), 3)
    s3.set_overrides(fl["ovr"][:, pick])
    s3.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"][:, pick])
    s3.advance(nsteps)
    np.testing.assert_array_equal(s3.get_states(), s.get_states()[pick])
    s.close()
    s3.close()


def test_per_vehicle_parameters_vs_oracle(golden18):
    """Vehicles given parameters of their own through set_vehicle_params, fields beyond the seven
    per-vehicle ones (CG height, track width, a shared wheelbase change, tire stiffness, steering limit),
    before the first launch and again after 5 s (which lays the slots out anew with the records kept):
    every vehicle within 1e-9 of the oracle run with its own parameter vector over 10 s; vehicles that
    were never edited are bit-identical to the same fleet without edits."""
    import lfdm_oracle as O

    say("[gpu18] per-vehicle parameter sets")
    flat = golden18["hmmwv_acc3/params"]
    n, half = 256, 5000
    fl = W.knot_fleet(n, flat, seed=12)
    F = W.FLAT18
    cx = W.tire_offsets(flat)[0] - 25 + 6

    def edited(idx, k, fac):
        f = flat.copy()
        f[idx] *= fac
        return k, f

    first = dict([edited(F["h"], 5, 1.15), edited(7, 17, 1.1), edited(cx, 100, 0.8), edited(F["maxSteer"], 200, 0.9)] +
                 [edited(F["a"], v, 1.03) for v in range(40, 48)])
    later = dict([edited(F["b"], 17, 0.97), edited(F["h"], 60, 0.9)])

    def solver():
        s = SolverHalfImplicitGPU(18, n)
        s.construct_flat(flat, n, None, TMeasy)
        s.Initialize(s.zero_record(), n)
        s.set_overrides(fl["ovr"])
        s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
        return s

    plain = solver()
    plain.advance(2 * half)
    base = plain.get_states()
    plain.close()
    s = solver()
    for v, f in first.items():
        s.set_vehicle_params(v, f)
    s.advance(half)
    mid = s.get_states()
    for v, f in later.items():
        s.set_vehicle_params(v, f)
    np.testing.assert_array_equal(s.get_states(), mid)  # the re-layout keeps every record
    s.advance(half)
    gpu = s.get_states()
    assert np.array_equal(s.get_vehicle_params(17), later[17]) and np.array_equal(s.get_vehicle_params(5), first[5])
    s.close()
    untouched = [v for v in range(n) if v not in first and v not in later]
    np.testing.assert_array_equal(gpu[untouched], base[untouched])
    worst = 0.0
    for v in sorted(set(first) | set(later)):
        f0 = first.get(v, flat)
        f1 = later.get(v, f0)
        k = fl["knots"][:, v:v + 1]
        ctrl = ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], k)
        o0 = fl["ovr"][:, v:v + 1] if v not in first else W.default_overrides(f0, 1)
        o1 = o0 if v not in later else W.default_overrides(f1, 1)
        m1, _ = O.run(18, f0, ctrl, 1, half, ovr=o0)
        # the second half starts at the time the solver reached by accumulating t += dt, as the oracle's loop
        m2, _ = O.run(18, f1, ctrl, 1, half, ovr=o1, init=m1, t0=W.time_grid(half + 1, 1e-3)[half])
        den = np.maximum(np.abs(m2[0]), 1.0)
        worst = max(worst, float((np.abs(gpu[v] - m2[0]) / den).max()))
    say(f"[gpu18] per-vehicle parameter sets: worst relative difference to the oracle {worst:.2e}")
    assert worst <= 1e-9


def test_chunking_and_block_size_invariance(golden18):
    """Results do not depend on kernel_sim_time (launch chunking) or threads per block."""
    say("[gpu18] chunking invariance")
    d = os.path.join(DATA, "18dof")
    res = []
    for ksim, tpb in ((2.0, 32), (10.0, 256), (0.5, 64)):
        s = SolverHalfImplicitGPU(18, 130)
        s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 130,
                    os.path.join(DATA, "input", "double_lane4.txt"))
        s.Initialize(s.zero_record(), 130)
        s.SetKernelSimTime(ksim)
        s.SetThreadsPerBlock(tpb)
        s.SetEndTime(10.0)
        s.Solve()
        res.append(s.get_states())
        s.close()
    np.testing.assert_array_equal(res[0], res[1])
    np.testing.assert_array_equal(res[0], res[2])


def test_solve_step_constant_controls(golden18):
    """SolveStep(t, steering, throttle, braking) with a controller-in-the-loop pattern against the
    oracle fed the same per-step controls (tests/18dof.cpp:113-165 uses the CPU step API)."""
    import lfdm_oracle as O

    say("[gpu18] SolveStep")
    flat = golden18["hmmwv_acc3/params"]
    n = 70
    s = SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.Initialize(s.zero_record(), n)
    s.SetKernelSimTime(0.1)
    t = 0.0
    ctrl = []
    while t < 5.0 - 1e-4:
        st, th = (0.0, 0.0) if t < 0.5 else (0.1, 0.5)
        t_new = s.SolveStep(t, st, th, 0.0)
        ctrl += [(st, th, 0.0)] * 100
        t = t_new
    gpu = s.get_states()
    c = np.repeat(np.array(ctrl)[:, :, None], n, axis=2)
    fin, _ = O.run18(flat, ("steps", c), n, c.shape[0])
    assert normwise_rel(gpu[None], fin[None]).max() <= TOL
    s.close()


def test_csv_output_format(tmp_path, golden18):
    """SetOutput writes <file>_<vehicle>.csv with the reference's header, zero first row and one row
    per output sample (dof18_halfImplicit_gpu.cu:342-399, 552-663)."""
    say("[gpu18] CSV output")
    d = os.path.join(DATA, "18dof")
    s = SolverHalfImplicitGPU(18, 100)
    s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 100, os.path.join(DATA, "input", "acc3.txt"))
    s.Initialize(s.zero_record(), 100)
    s.SetEndTime(10.0)
    base = str(tmp_path / "run")
    s.SetOutput(base, 10, False, 5)
    s.Solve()
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("run_"))
    assert len(files) == 5
    text = open(os.path.join(tmp_path, files[0])).read().splitlines()
    assert text[0] == "time,x,y,vx,vy,roll,yaw,roll_rate,yaw_rate,wlf,wrf,wlr,wrr,"
    assert text[1] == ",".join(["0"] * 13) + ","
    assert len(text) == 2 + 100
    last = [float(v) for v in text[-1].rstrip(",").split(",")]
    assert abs(last[0] - 10.0) < 1e-6
    ref = golden18["hmmwv_acc3/final"][0]
    assert abs(last[1] - ref[60]) <= 1e-4 * abs(ref[60])
    s.close()
