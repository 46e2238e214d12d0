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

import numpy as np
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



JAC = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "jacobian_check")
JAC_CASES = ["hmmwv_t1", "hmmwv_t4", "hmmwv_t8", "sedan_rwd_t3", "audi_t5"]
DT = 1e-3


def our_step_jacobian(veh, tire, rec, ctrl, dt):
    """(u', v', d(step)/d(packed state), d(step)/d(steering, throttle)) from the step adjoint."""
    d = os.path.join(DATA, "18dof")
    inp = " ".join(repr(float(x)) for x in (*ctrl, dt, *rec))
    r = subprocess.run([JAC, f"{d}/{veh}", f"{d}/{tire}"], input=inp, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    rows = [[float(x) for x in line.split()] for line in r.stdout.splitlines()]
    m = np.array(rows[1:])
    return rows[0][0], rows[0][1], m[:, :-2], m[:, -2:]


def packed_fields(n):
    """Record fields of the reference's packY order (TMeasy): xe, ye per tire, omega x 4, crank with
    a converter, then x y u v psi wz phi wx."""
    f = [15 * i + k for i in range(4) for k in (0, 1)] + [15 * i + 4 for i in range(4)]
    return f + ([72] if n == 21 else []) + list(range(60, 68))


def fd_columns(flat, base, ctrl, fields, hs, hc):
    """d(one oracle step)/d(packed state, steering, throttle) by central differences with state steps hs
    and control steps hc; rows and columns in packY order."""
    import lfdm_oracle as O

    m = len(fields)
    recs = np.repeat(base[None], 2 * (m + 2), axis=0)
    c = np.repeat(np.asarray(ctrl, dtype=np.float64)[None], 2 * (m + 2), axis=0)
    for j, f in enumerate(fields):
        recs[2 * j, f] += hs[j]
        recs[2 * j + 1, f] -= hs[j]
    for j in range(2):
        c[2 * (m + j), j] += hc[j]
        c[2 * (m + j) + 1, j] -= hc[j]
    fin, _ = O.run18(flat, ("steps", c.T[None].copy()), recs.shape[0], 1, init=recs)
    out = fin[:, fields]
    h = np.concatenate([hs, hc])
    return ((out[0::2] - out[1::2]) / (2 * h[:, None])).T


@pytest.mark.parametrize("case", JAC_CASES)
def test_step_jacobian_matches_reference_integrate_step_with_jacobian(case):
    """The step adjoint against the reference's own IntegrateStepWithJacobian
    (wheeled_vehicle_models/18dof/dof18_halfImplicit.cpp:356-463; fixtures from
    tests/golden/gen_jacobian18.py): central differences of the right-hand side f at gear 0 with zero
    lagged accelerations (PerturbRhsFun, :679-745), state step 1e-3 |y| (1e-3 where |y| < 1e-5),
    control step 1e-3, braking 0.

    One half-implicit step maps the packed state y to y + dt f(y) for the tire deflections, wheel and
    crank speeds and the velocities; psi and phi advance with the new wz and wx, x and y with the new
    u and v at the old psi (:483-498).  So the reference's f-Jacobian fixes the whole step Jacobian.
    Three checks, all in right-hand-side units ((step Jacobian - I) / dt), row by row relative to the
    row's largest entry:
    1. the oracle's step, differenced with the reference's own steps, reproduces the reference's
       Jacobian to 1e-8 (so both difference the same function; the oracle differences the stepped
       state y + dt f, whose rounding adds up to ulp(y) / (2 h dt) at a step h);
    2. the adjoint (seeded with each packed output in turn) equals Richardson-extrapolated central
       differences of the oracle's step (steps 1e-5 max(|y|, 1) and 1e-5 for the controls) to 1e-6;
    3. the adjoint against the reference's Jacobian directly: within 2e-2.  The reference's steps
       are large: at hmmwv_t4 the LF wheel speed's stencil (+-0.015 rad/s) straddles a kink of the
       differential split, and for audi_t5 the u column carries 1.7e-2 of second-order truncation
       (it falls a hundredfold per tenfold smaller step)."""
    import sys

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "low-fidelity-dynamic-models_amd"))
    import lfdm_amd

    assert os.path.exists(JAC), "run make -C low-fidelity-dynamic-models_amd adjcheck"
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "d18_jacobian.npz"))
    veh, tire = (str(x) for x in g[f"{case}/files"])
    rec, ctrl = g[f"{case}/record"], g[f"{case}/controls"]
    J, C = g[f"{case}/jac_state"], g[f"{case}/jac_controls"]
    n = J.shape[0]
    up, vp, D, E = our_step_jacobian(veh, tire, rec, ctrl, DT)
    assert D.shape == (n, n)
    eye = np.hstack([np.eye(n), np.zeros((n, 2))])
    ours = (np.hstack([D, E]) - eye) / DT

    def rowwise(a, b):
        return np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-12)

    # the reference's step rule, and the step started as PerturbRhsFun starts
    fields = packed_fields(n)
    hs = np.abs(rec[fields] * 1e-3)
    hs[hs < 1e-8] = 1e-3
    hc = np.array([1e-3, 1e-3])
    base = rec.copy()
    for i in range(4):
        base[15 * i + 2] = base[15 * i + 3] = 0.0
    base[68] = base[69] = base[74] = 0.0
    flat = lfdm_amd.params_from_json(18, os.path.join(DATA, "18dof", veh), os.path.join(DATA, "18dof", tire))

    # 1. the rows the step advances by dt f (all but x, y, psi, phi)
    dyn = list(range(n - 8)) + [n - 6, n - 5, n - 3, n - 1]
    fd_ref = (fd_columns(flat, base, ctrl, fields, hs, hc) - eye) / DT
    e1 = rowwise(fd_ref[dyn], np.hstack([J, C])[dyn])
    print(f"1. oracle differences with the reference's steps vs its Jacobian: worst row {e1.max():.2e}")
    assert e1.max() <= 1e-8

    # 2. Richardson extrapolation of central differences with steps 1e-5 max(|y|, 1) and twice that
    h2 = 1e-5 * np.maximum(np.abs(rec[fields]), 1.0)
    f1 = fd_columns(flat, base, ctrl, fields, h2, hc / 100)
    f2 = fd_columns(flat, base, ctrl, fields, 2 * h2, hc / 50)
    fd = ((4 * f1 - f2) / 3 - eye) / DT
    e2 = rowwise(ours, fd)
    print(f"2. adjoint vs extrapolated oracle differences: worst row {e2.max():.2e}")
    assert e2.max() <= 1e-6

    # 3. directly against the reference: the step Jacobian its f-Jacobian implies
    ix = {k: n - 8 + i for i, k in enumerate(("x", "y", "u", "v", "psi", "wz", "phi", "wx"))}
    P = np.hstack([np.eye(n) + DT * J, DT * C])
    for pos, rate in (("psi", "wz"), ("phi", "wx")):
        P[ix[pos]] = eye[ix[pos]] + DT * P[ix[rate]]
    c, s = np.cos(rec[64]), np.sin(rec[64])
    Pu, Pv = P[ix["u"]] - eye[ix["u"]], P[ix["v"]] - eye[ix["v"]]
    for pos, a, b, dpsi in (("x", c, -s, -up * s - vp * c), ("y", s, c, up * c - vp * s)):
        P[ix[pos]] = eye[ix[pos]] + DT * (a * (eye[ix["u"]] + Pu) + b * (eye[ix["v"]] + Pv))
        P[ix[pos], ix["psi"]] += DT * dpsi
    e3 = rowwise(ours, (P - eye) / DT)
    print(f"3. adjoint vs the reference's Jacobian: worst row {e3.max():.2e}, median {np.median(e3):.2e}")
    assert e3.max() <= 2e-2
