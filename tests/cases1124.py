"""The 11- and 24-DOF golden cases (tests/golden/gen_golden1124.py) as inputs for the oracle and
the HIP solver."""
import os

import numpy as np

from conftest import DATA
from lfdm_amd import workloads as W

INP = os.path.join(DATA, "input")

# dof -> name -> (vehicle, tire, TMeasyNr?, control kind, driver file or None)
CASES = {
    11: {
        "hmmwv_acc3": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
        "hmmwv_step_api": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "step_api", None),
        "sedan_nr_acc05": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "file", "acc05.txt"),
        "hmmwv_acc3_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
        "hmmwv_dl4_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "double_lane4.txt"),
        "sedan_rwd_acc3": ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", False, "file", "acc3.txt"),
        "audi_dl4": ("audi/vehicle.json", "audi/tmeasy.json", False, "file", "double_lane4.txt"),
        "art_knots": ("ART/vehicle.json", "ART/tmeasy.json", False, "knots", None),
        "hmmwv_knot_fleet": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "knots", None),
        "sedan_nr_knots": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "knots", None),
    },
    24: {
        "hmmwv_acc3": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
        "hmmwv_step_api": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "step_api", None),
        "sedan_nr_acc05": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "file", "acc05.txt"),
        "hmmwv_acc3_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "acc3.txt"),
        "hmmwv_dl4_22s": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "file", "double_lane4.txt"),
        "sedan_dl4": ("Sedan/vehicle.json", "Sedan/tmeasy.json", False, "file", "double_lane4.txt"),
        "hmmwv_knot_fleet": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "knots", None),
        "sedan_nr_knots_u10": ("Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "knots", None),
        "tall_offground": ("../custom/tall_vehicle24.json", "HMMWV/tmeasy.json", False, "knots", None),
    },
}

# Reference known-answer values: final x within +-1 m (wheeled_vehicle_models/tests/11dof.cpp:36,156,239,
# tests/11dof_gpu.cu:76-77, tests/24dof.cpp:43,176,273, tests/24dof_gpu.cu:104-105).
KAT_X = {
    11: {"hmmwv_acc3": 70.6758, "hmmwv_step_api": 57.3525, "sedan_nr_acc05": 115.1, "hmmwv_acc3_22s": 198.72,
         "hmmwv_dl4_22s": 104.805},
    24: {"hmmwv_acc3": 89.1625, "hmmwv_step_api": 60.6147, "sedan_nr_acc05": 114.938, "hmmwv_acc3_22s": 235.685,
         "hmmwv_dl4_22s": 91.5211},
}
# index of the vehicle x position in a record (from the end)
X_FROM_END = {11: -12, 18: -15, 24: -24}


def suspension_path(dof, veh):
    if dof != 24:
        return None
    top = veh.split("/")[0]
    return os.path.join(DATA, "24dof", "HMMWV" if top == ".." else top, "suspension.json")


def paths(dof, name):
    veh, tire, nr, kind, drv = CASES[dof][name]
    d = os.path.join(DATA, f"{dof}dof")
    return (os.path.join(d, veh), os.path.join(d, tire), nr, kind, os.path.join(INP, drv) if drv else None,
            suspension_path(dof, veh))


def step_api_controls(nsteps, dt=1e-3):
    t = W.time_grid(nsteps, dt)
    c = np.zeros((nsteps, 3, 1))
    on = t > 0.5
    c[on, 0, 0] = 0.1
    c[on, 1, 0] = 0.5
    return c


def inputs(dof, name, g):
    """(flat params, controls for the oracle, n, ovr or None, initial records (n, rec)) of one case."""
    veh, tire, nr, kind, drv, _ = paths(dof, name)
    flat = g[f"{name}/params"]
    nsteps = int(g[f"{name}/spec"][0])
    init = g[f"{name}/init"]
    if kind == "file":
        rows = np.loadtxt(drv, ndmin=2)[:, :4]
        return flat, ("table", rows), 1, None, init
    if kind == "step_api":
        return flat, ("steps", step_api_controls(nsteps)), 1, None, init
    if kind == "knots":
        ns, ds, np_, dp = g[f"{name}/knot_meta"]
        knots = g[f"{name}/knots"]
        return (flat, ("knots", int(ns), float(ds), int(np_), float(dp), knots), knots.shape[1], g[f"{name}/ovr"],
                init)
    raise ValueError(kind)
