"""Pins the 11- and 24-DOF CPU oracle (oracle/orc11.h, oracle/orc24.h) to the reference.

* golden trajectories written by the reference solvers themselves (tests/golden/gen_golden1124.py):
  bit-identical on every case (driver files, step API, knot programs with per-vehicle parameters,
  a 24-DOF vehicle that lifts wheels, a 24-DOF start at 10 m/s);
* the reference's own known-answer values (tests/11dof*.c*, tests/24dof*.c*);
* the 24-DOF Initialize (initializeTireSus, dof24.cpp:34-89) against the reference's initial records
  and against static equilibrium (the four tire loads carry the whole weight);
* live against oracle/_ref (the reference compiled from its sources) where it is built;
* invariants that a wrong sign, index or dropped term breaks: rest stays rest, straight-line
  driving keeps every lateral state at exactly zero, mirrored steering mirrors the trajectory,
  and a vehicle's result does not depend on the batch it runs in.
"""
import numpy as np
import pytest

import lfdm_oracle as O
import refbind
from cases1124 import CASES, KAT_X, X_FROM_END, inputs, paths

G = 9.81
CASE_IDS = [(dof, name) for dof in (11, 24) for name in sorted(CASES[dof])]
# vehicle block of a record: 11-DOF x y u v psi wz ...; 24-DOF x y z u v w psi wz phi wx theta wy ...
VEH = {11: 12, 24: 24}
LATERAL = {11: (1, 3, 4, 5, 7, 8), 24: (1, 4, 6, 7, 8, 9, 13, 15, 17, 19, 20)}  # y v psi wz (phi wx) + rates


def golden(request, dof):
    return request.getfixturevalue(f"golden{dof}")


def run_case(dof, name, g, nsteps=None):
    flat, ctrl, n, ovr, init = inputs(dof, name, g)
    spec = g[f"{name}/spec"]
    steps, every, nr = int(spec[0]), int(spec[1]), bool(spec[2])
    return O.run(dof, flat, ctrl, n, nsteps or steps, tire_nr=nr, ovr=ovr, init=init,
                 every=0 if nsteps else every)


@pytest.mark.parametrize("dof,name", CASE_IDS)
def test_oracle_matches_golden(dof, name, request):
    g = golden(request, dof)
    fin, traj = run_case(dof, name, g)
    np.testing.assert_array_equal(traj, g[f"{name}/traj"])
    np.testing.assert_array_equal(fin, g[f"{name}/final"])


@pytest.mark.parametrize("dof,name", [(d, n) for d in (11, 24) for n in sorted(KAT_X[d])])
def test_reference_known_answers(dof, name, request):
    fin, _ = run_case(dof, name, golden(request, dof))
    assert abs(fin[0, X_FROM_END[dof]] - KAT_X[dof][name]) < 1.0


@pytest.mark.parametrize("name", sorted(CASES[24]))
def test_init24_matches_reference_initialize(name, golden24):
    """Zero records (+ the case's initial speed) through initializeTireSus reproduce the records the
    reference's Initialize produced, bit for bit, with per-vehicle parameters where the case has them."""
    flat, _, n, ovr, init = inputs(24, name, golden24)
    nr = bool(golden24[f"{name}/spec"][2])
    u0 = golden24[f"{name}/spec"][3]
    rec = np.zeros((n, O.record_len(24, nr)))
    rec[:, -24 + 3] = u0
    np.testing.assert_array_equal(O.init24(flat, rec, nr, ovr), init)


def test_init24_static_equilibrium(golden24):
    """At rest the four tire loads carry the whole vehicle: sum fz = g (m + 2 muf + 2 mur)."""
    for name in ("hmmwv_acc3", "sedan_nr_acc05"):
        flat, _, _, _, init = inputs(24, name, golden24)
        nr = bool(golden24[f"{name}/spec"][2])
        tl = 19 if nr else 23
        fz = [init[0, k * tl + (7 if nr else 11)] for k in range(4)]
        m, muf, mur = flat[3], flat[9], flat[10]
        assert abs(sum(fz) - G * (m + 2 * muf + 2 * mur)) <= 1e-9 * sum(fz)
        assert fz[0] == fz[1] and fz[2] == fz[3]


def test_static_loads_11(golden11):
    """11-DOF: after one step at rest the axle loads are Z1 and Z3 of dof11.cpp:237-244, which sum to
    g (m + muf + mur) / 2 (each axle tire carries the per-tire load, its parameters are doubled)."""
    flat, _, _, _, init = inputs(11, "hmmwv_acc3", golden11)
    fin, _ = O.run(11, flat, ("steps", np.zeros((1, 3, 1))), 1, 1, init=init)
    a, b, m, muf, mur = flat[0], flat[1], flat[3], flat[5], flat[6]
    fz0, fz1 = fin[0, 10], fin[0, 15 + 10]
    assert abs(fz0 - (m * G * b / (2 * (a + b)) + muf * G / 2)) <= 1e-12 * fz0
    assert abs(fz1 - (m * G * a / (2 * (a + b)) + mur * G / 2)) <= 1e-12 * fz1


@pytest.mark.skipif(not (refbind.available(11) and refbind.available(24)),
                    reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_live_reference_random_controls(seed):
    rng = np.random.default_rng(seed)
    n = 3000
    t = np.arange(n) * 1e-3
    ctrl = np.stack([0.6 * np.sin(t * rng.uniform(0.5, 3)) * (t > 0.3), np.clip(t / rng.uniform(0.5, 2), 0, 1),
                     (rng.random(n) < 0.05) * rng.uniform(0, 1)], 1)
    for dof, name in ((11, "hmmwv_acc3"), (11, "sedan_nr_acc05"), (11, "art_knots"), (24, "hmmwv_acc3"),
                      (24, "sedan_nr_acc05"), (24, "sedan_dl4")):
        veh, tire, nr, _, _, sus = paths(dof, name)
        r = refbind.RefSolver(dof, veh, tire, nr, sus=sus, u0=3.0 * seed)
        init = r.record()
        _, ref = r.run(n, ctrl, every=7)
        _, out = O.run(dof, r.params(), ("steps", ctrl[:, :, None]), 1, n, tire_nr=nr, init=init[None], every=7)
        np.testing.assert_array_equal(out[:, 0, :], ref)


@pytest.mark.parametrize("dof", [11, 24])
def test_rest_straight_and_mirror_invariants(dof, request):
    g = golden(request, dof)
    flat, _, _, _, init = inputs(dof, "hmmwv_acc3", g)
    n = 3000
    vb = O.record_len(dof) - VEH[dof]
    # zero controls from the initialized rest state: the chassis never moves
    fin, _ = O.run(dof, flat, ("steps", np.zeros((n, 3, 1))), 1, n, init=init)
    assert np.all(fin[0, vb:vb + VEH[dof] - 3] == 0.0)
    # straight-line throttle: every lateral state stays exactly zero while the vehicle accelerates
    c = np.zeros((n, 3, 1))
    c[:, 1] = 0.6
    fin, _ = O.run(dof, flat, ("steps", c), 1, n, init=init)
    assert fin[0, vb + (2 if dof == 11 else 3)] > 1.0
    assert np.all(fin[0, [vb + i for i in LATERAL[dof]]] == 0.0)
    # mirrored steering mirrors the trajectory: lateral states flip sign, the others are unchanged
    t = np.arange(n) * 1e-3
    c[:, 0, 0] = 0.3 * np.sin(2 * t)
    f1, _ = O.run(dof, flat, ("steps", c), 1, n, init=init)
    c[:, 0, 0] *= -1
    f2, _ = O.run(dof, flat, ("steps", c), 1, n, init=init)
    sign = np.ones(VEH[dof])
    sign[list(LATERAL[dof])] = -1
    a, b = f1[0, vb:], f2[0, vb:] * sign
    assert np.abs(f1[0, vb + 1]) > 0.1  # the manoeuvre does move sideways
    np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("dof", [11, 24])
def test_batch_independence(dof, request):
    g = golden(request, dof)
    flat, ctrl, n, ovr, init = inputs(dof, "hmmwv_knot_fleet", g)
    fin_all, _ = O.run(dof, flat, ctrl, n, 3000, ovr=ovr, init=init)
    _, ns, ds, np_, dp, knots = ctrl
    for i in (0, 5, 15):
        one = ("knots", ns, ds, np_, dp, knots[:, i:i + 1])
        fin1, _ = O.run(dof, flat, one, 1, 3000, ovr=ovr[:, i:i + 1], init=init[i:i + 1])
        np.testing.assert_array_equal(fin1[0], fin_all[i])
