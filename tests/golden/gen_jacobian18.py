"""Writes the 18-DOF right-hand-side Jacobian fixtures from the reference CPU solver.

Run in a container that holds the reference sources, after `make -C oracle ref`:
    python tests/golden/gen_jacobian18.py
For each case it drives oracle/_ref/libref18.so (the reference's wheeled_vehicle_models/18dof sources
plus oracle/ref_shim18.cpp) with IntegrateStep to an operating point, stores the full state record
there, then calls d18SolverHalfImplicit::IntegrateStepWithJacobian
(wheeled_vehicle_models/18dof/dof18_halfImplicit.cpp:356-463) and stores its central-difference
Jacobians of the right-hand side with respect to the packed states (packY order, :320-341) and to
(steering, throttle).  The reference evaluates that right-hand side on freshly constructed states
(PerturbRhsFun, :679-745): gear 0, lagged accelerations 0.  Output: tests/golden/d18_jacobian.npz.
Only data is written; no reference source is copied.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import refbind  # noqa: E402

D18 = os.path.join(HERE, "data", "18dof")
DT = 1e-3

# name: (vehicle json, tire json, seconds to the operating point, steering amplitude, throttle)
CASES = {
    "hmmwv_t1": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", 1.0, 0.25, 0.6),
    "hmmwv_t4": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", 4.0, 0.25, 0.6),
    "hmmwv_t8": ("HMMWV/vehicle.json", "HMMWV/tmeasy.json", 8.0, 0.4, 0.9),
    "sedan_rwd_t3": ("Sedan/vehicle_RWD.json", "Sedan/tmeasy.json", 3.0, 0.2, 0.5),
    "audi_t5": ("audi/vehicle.json", "audi/tmeasy.json", 5.0, 0.3, 0.7),
}


def controls(nsteps, amp, thr):
    """Steering amp sin(2 pi 0.2 t) + 0.05, throttle thr min(t / 2, 1), no braking (the reference's
    control Jacobian perturbs with braking 0)."""
    t = np.arange(nsteps) * DT
    c = np.zeros((nsteps, 3))
    c[:, 0] = amp * np.sin(2 * np.pi * 0.2 * t) + 0.05
    c[:, 1] = thr * np.minimum(t / 2.0, 1.0)
    return c


def main():
    assert refbind.available(18), "build oracle/_ref first: make -C oracle ref"
    out = {}
    names = []
    for name, (veh, tire, secs, amp, thr) in CASES.items():
        n = int(round(secs / DT))
        c = controls(n + 1, amp, thr)
        r = refbind.Ref18(os.path.join(D18, veh), os.path.join(D18, tire), False, None, DT)
        t, _ = r.run(n, c[:n])
        rec = r.record()
        js, jc = r.step_jacobian(t, c[n, 0], c[n, 1], c[n, 2])
        r.close()
        out[f"{name}/record"] = rec
        out[f"{name}/controls"] = c[n]
        out[f"{name}/jac_state"] = js
        out[f"{name}/jac_controls"] = jc
        out[f"{name}/files"] = np.array([veh, tire])
        names.append(name)
        print(f"{name}: t = {t:.3f}, n = {js.shape[0]}, gear {int(rec[-1])}, u = {rec[62]:.3f}")
    out["cases"] = np.array(names)
    path = os.path.join(HERE, "d18_jacobian.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
