"""The 18-DOF golden cases (tests/golden/gen_golden.py) as inputs for the oracle and the HIP solver."""
import os

import numpy as np

from conftest import DATA
from lfdm_amd import workloads as W

D18 = os.path.join(DATA, "18dof")
INP = os.path.join(DATA, "input")

# name -> (vehicle, tire, TMeasyNr?, control kind, driver file or None)
CASES = {
    "hmmwv_acc3": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
    "hmmwv_acc3_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
    "hmmwv_dl4_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "double_lane4.txt"),
    "hmmwv_step_api": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "step_api", None),
    "sedan_nr_acc05": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "file", "acc05.txt"),
    "sedan_nr_knots": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "knots", None),
    "sedan_rwd_acc3": ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", False, "file", "acc3.txt"),
    "audi_dl4": ("audi/vehicle.json", "audi/tmeasy.json", False, "file", "double_lane4.txt"),
    "art_knots": ("ART/vehicle.json", "ART/tmeasy.json", False, "knots", None),
    "hmmwv_knot_fleet": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "knots", None),
    "hmmwv_ramp_sine": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "ramp_sine", None),
    "tall_offground": ("../custom/tall_vehicle.json", "HMMWV/tmeasy.json", False, "knots", None),
}

# Reference known-answer values: final x within +-1 m (wheeled_vehicle_models/tests/18dof.cpp:38,162,249,
# tests/18dof_gpu.cu:52,93-94).
KAT_X = {"hmmwv_acc3": 88.5335, "hmmwv_step_api": 62.7973, "sedan_nr_acc05": 114.973, "hmmwv_acc3_22s": 241.112,
         "hmmwv_dl4_22s": 106.548}


def paths(name):
    veh, tire, nr, kind, drv = CASES[name]
    return (os.path.join(D18, veh), os.path.join(D18, tire), nr, kind,
            os.path.join(INP, drv) if drv else None)


def step_api_controls(nsteps, dt=1e-3):
    t = W.time_grid(nsteps, dt)
    c = np.zeros((nsteps, 3, 1))
    on = t > 0.5
    c[on, 0, 0] = 0.1
    c[on, 1, 0] = 0.5
    return c


def inputs(name, g):
    """(flat params, controls description for the oracle, n, ovr or None) of one case."""
    veh, tire, nr, kind, drv = paths(name)
    flat = g[f"{name}/params"]
    nsteps = int(g[f"{name}/spec"][0])
    if kind == "file":
        rows = np.loadtxt(drv, ndmin=2)[:, :4]
        return flat, ("table", rows), 1, None
    if kind == "step_api":
        return flat, ("steps", step_api_controls(nsteps)), 1, None
    if kind == "knots":
        ns, ds, np_, dp = g[f"{name}/knot_meta"]
        knots = g[f"{name}/knots"]
        return flat, ("knots", int(ns), float(ds), int(np_), float(dp), knots), knots.shape[1], g[f"{name}/ovr"]
    if kind == "ramp_sine":
        seed, n = (int(v) for v in g[f"{name}/seed"])
        ctrl = W.ramp_sine(n, nsteps=nsteps, seed=seed, max_steer=flat[W.FLAT18["maxSteer"]])
        return flat, ("steps", ctrl), n, None
    raise ValueError(kind)


def normwise_rel(a, b, axis=0):
    """max_t |a - b| / max_t |b| per (vehicle, channel); 0/0 counts as 0."""
    num = np.abs(a - b).max(axis=axis)
    den = np.abs(b).max(axis=axis)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


def final_rel(fin, traj_ref, fin_ref):
    """Normwise error of final states, scaled like the trajectories: |a - b| / max_t |b| with the
    maximum taken over the reference trajectory and its final state."""
    den = np.maximum(np.abs(traj_ref).max(axis=0), np.abs(fin_ref))
    num = np.abs(fin - fin_ref)
    return np.where(den > 0, num / np.where(den > 0, den, 1), num)


ULP_FIELDS = (0, 1, 3)  # m, jz, dfy0Pn: per-vehicle parameters nudged by one ulp


def ulp_spread(O, flat, ctrl, n, nsteps, ovr, every, base_traj=None):
    """The reference model's own sensitivity: max normwise change of each vehicle's sampled
    trajectory when one of its parameters moves by one ulp (up or down), per vehicle."""
    if base_traj is None:
        _, base_traj = O.run18(flat, ctrl, n, nsteps, ovr=ovr, every=every)
    spread = np.zeros(n)
    for f in ULP_FIELDS:
        for direction in (np.inf, -np.inf):
            o2 = ovr.copy()
            o2[f] = np.nextafter(o2[f], direction)
            _, tr = O.run18(flat, ctrl, n, nsteps, ovr=o2, every=every)
            spread = np.maximum(spread, normwise_rel(tr, base_traj).max(axis=-1))
    return spread
