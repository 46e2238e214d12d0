"""Pins the CPU oracle (oracle/orc18.h) to the reference.

* golden trajectories written by the reference solver itself (tests/golden/gen_golden.py):
  bit-identical on every case with inputs that are exactly reproducible (driver files, step API,
  knot programs), within 1e-12 normwise on the ramp_sine case whose controls come from numpy's sin;
* the reference's own known-answer values;
* live against oracle/_ref (the reference compiled from its sources) where it is built;
* closed forms and invariants of the building blocks (chosen so a dropped term, a wrong sign or
  a swapped operand fails one of them).
"""
import numpy as np
import pytest

import lfdm_oracle as O
import refbind
from cases18 import CASES, KAT_X, inputs, normwise_rel, paths
from conftest import DATA

# record fields of a TMeasy vehicle that the reference leaves uninitialized when the vehicle has no
# torque converter (VehicleState has no initializer for _dOmega_crank, dof18.h:193-210)
DCRANK = -2


def run_case(name, g):
    flat, ctrl, n, ovr = inputs(name, g)
    nsteps, every, nr = (int(v) for v in g[f"{name}/spec"])
    fin, traj = O.run18(flat, ctrl, n, nsteps, tire_nr=bool(nr), ovr=ovr, every=every)
    return fin, traj


def comparable(name, g, a):
    a = a.copy()
    flat = g[f"{name}/params"]
    if flat[21] == 0:  # tcbool
        a[..., DCRANK] = 0.0
    return a


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_matches_golden(name, golden18):
    fin, traj = run_case(name, golden18)
    ref_traj = comparable(name, golden18, golden18[f"{name}/traj"])
    ref_fin = comparable(name, golden18, golden18[f"{name}/final"])
    traj = comparable(name, golden18, traj)
    fin = comparable(name, golden18, fin)
    if CASES[name][3] == "ramp_sine":
        assert normwise_rel(traj, ref_traj).max() <= 1e-12
        assert normwise_rel(fin[None], ref_fin[None]).max() <= 1e-12
    else:
        np.testing.assert_array_equal(traj, ref_traj)
        np.testing.assert_array_equal(fin, ref_fin)


@pytest.mark.parametrize("name", sorted(KAT_X))
def test_reference_known_answers(name, golden18):
    fin, _ = run_case(name, golden18)
    assert abs(fin[0, -15] - KAT_X[name]) < 1.0


@pytest.mark.skipif(not refbind.available(18), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_live_reference_random_controls(seed, golden18):
    rng = np.random.default_rng(seed)
    n = 3000
    t = np.arange(n) * 1e-3
    ctrl = np.stack([0.6 * np.sin(t * rng.uniform(0.5, 3)) * (t > 0.3), np.clip(t / rng.uniform(0.5, 2), 0, 1),
                     (rng.random(n) < 0.05) * rng.uniform(0, 1)], 1)
    for name in ("hmmwv_acc3", "sedan_nr_acc05", "art_knots", "sedan_rwd_acc3"):
        veh, tire, nr, _, _ = paths(name)
        r = refbind.Ref18(veh, tire, nr)
        _, ref = r.run(n, ctrl, every=7)
        flat = r.params()
        _, out = O.run18(flat, ("steps", ctrl[:, :, None]), 1, n, tire_nr=nr, every=7)
        out = out[:, 0, :]
        if flat[21] == 0:
            out[:, DCRANK] = 0
            ref[:, DCRANK] = 0
        np.testing.assert_array_equal(out, ref)


def test_interp_closed_forms():
    # InterpL / InterpQ reproduce w1 at pn and w2 at 2 pn (utils.h:93-100)
    for w1, w2, pn in [(3.0, 5.0, 100.0), (151447.29, 236412.79, 8562.8266), (-2.0, 7.5, 3.0)]:
        assert O.interp_l(pn, w1, w2, pn) == w1
        assert abs(O.interp_l(2 * pn, w1, w2, pn) - w2) <= 1e-12 * abs(w2)
        assert abs(O.interp_q(pn, w1, w2, pn) - w1) <= 1e-12 * abs(w1)
        assert abs(O.interp_q(2 * pn, w1, w2, pn) - w2) <= 1e-12 * abs(w2)
        assert O.interp_q(0.0, w1, w2, pn) == 0.0


def test_tmxy_closed_forms():
    df0, sm, fm, ss, fs = 50000.0, 0.15, 6000.0, 0.9, 5000.0
    # peak at sm, sliding value beyond ss, zero at zero slip, fos = f / s
    f, fos = O.tmxy(sm, df0, sm, fm, ss, fs)
    assert abs(f - fm) <= 1e-9 * fm
    f, fos = O.tmxy(1.2, df0, sm, fm, ss, fs)
    assert f == fs and abs(fos - fs / 1.2) <= 1e-15 * fos
    assert O.tmxy(0.0, df0, sm, fm, ss, fs) == (0.0, 0.0)
    # initial slope: f ~ max(2 fm / sm, df0) * s for tiny s
    s = 1e-7
    f, fos = O.tmxy(s, df0, sm, fm, ss, fs)
    assert abs(fos - max(2 * fm / sm, df0)) <= 1e-3 * fos
    # continuity at ss and across the parabola junction
    a, _ = O.tmxy(ss * (1 - 1e-12), df0, sm, fm, ss, fs)
    assert abs(a - fs) <= 1e-6 * fs
    xs = np.linspace(sm, ss, 2001)
    ys = np.array([O.tmxy(x, df0, sm, fm, ss, fs)[0] for x in xs])
    assert np.all(np.abs(np.diff(ys)) < 10.0)


def test_sine_step_and_map():
    assert O.sine_step(0.5, 1, 1, 1.5, 0) == 1
    assert O.sine_step(2.0, 1, 1, 1.5, 0) == 0
    assert abs(O.sine_step(1.25, 1, 1, 1.5, 0) - 0.5) < 1e-15
    m = np.array([[0.0, 2.0], [1.0, 4.0], [3.0, 0.0]])
    assert O.map_y(m, -1.0) == 2.0 and O.map_y(m, 5.0) == 0.0
    assert O.map_y(m, 1.0) == 4.0 and O.map_y(m, 0.5) == 3.0 and O.map_y(m, 2.0) == 2.0


def test_differential_split():
    # torque conserved; slower wheel takes the larger share once the speed gap exceeds 0.25 rad/s
    for tq, sl, sr in [(100.0, 10.0, 10.1), (100.0, 10.0, 10.4), (-50.0, 3.0, 9.0), (80.0, 9.0, 3.0)]:
        a, b = O.diff_split(tq, sl, sr)
        assert abs(a + b - tq) <= 1e-12 * abs(tq)
    a, b = O.diff_split(90.0, 1.0, 2.0)  # bias 2 -> slow wheel gets 2/3
    assert abs(a - 60.0) < 1e-12 and abs(b - 30.0) < 1e-12
    a, b = O.diff_split(90.0, 2.0, 1.0)
    assert abs(b - 60.0) < 1e-12


def test_static_loads_and_chassis_solve(golden18):
    flat = golden18["hmmwv_acc3/params"]
    rl = 75
    rec = np.zeros(rl)
    loads = O.loads18(flat, rec)
    m, muf, mur = flat[3], flat[9], flat[10]
    # at rest the loads carry the sprung mass and half of each axle's unsprung mass (as coded)
    assert abs(loads.sum() - 9.81 * (m + muf + mur)) <= 1e-9 * loads.sum()
    assert loads[0] == loads[1] and loads[2] == loads[3]
    # the closed-form 3x3 chassis solve satisfies its own linear system (dof18.cpp:1019-1078)
    rng = np.random.default_rng(3)
    a, b, h, jz, jx, jxz, cf, cr = flat[0], flat[1], flat[2], flat[4], flat[5], flat[6], flat[7], flat[8]
    hrcf, hrcr, krof, kror, brof, bror = flat[11], flat[12], flat[13], flat[14], flat[15], flat[16]
    for _ in range(5):
        rec = np.zeros(rl)
        u, v, wz, wx, phi = rng.uniform(-5, 20), rng.uniform(-1, 1), rng.uniform(-0.5, 0.5), rng.uniform(-.2, .2), \
            rng.uniform(-0.05, 0.05)
        rec[60 + 2], rec[60 + 3], rec[60 + 5], rec[60 + 7], rec[60 + 6] = u, v, wz, wx, phi
        fx, fy = rng.uniform(-3000, 3000, 4), rng.uniform(-3000, 3000, 4)
        udot, vdot, wxdot, wzdot = O.veh_rhs18(flat, rec, fx, fy)
        mt = m + 2 * (muf + mur)
        hrc = (hrcf * b + hrcr * a) / (a + b)
        A1, A2, A3 = mur * b - muf * a, jx + m * hrc ** 2, hrc * m
        E1 = -mt * wz * u + fy.sum()
        E2 = (fy[0] + fy[1]) * a - (fy[2] + fy[3]) * b + (fx[1] - fx[0]) * cf / 2 + (fx[3] - fx[2]) * cr / 2 + (
            -muf * a + mur * b) * wz * u
        E3 = m * 9.81 * hrc * phi - (krof + kror) * phi - (brof + bror) * wx + hrc * m * wz * u
        M = np.array([[mt, -A1, -A3], [-A1, jz, jxz], [-A3, jxz, A2]])  # symmetric mass matrix
        x = np.array([vdot, wzdot, wxdot])
        r = M @ x - np.array([E1, E2, E3])
        assert np.abs(r).max() <= 1e-9 * np.abs([E1, E2, E3]).max()


def test_rest_and_symmetry_invariants(golden18):
    flat = golden18["hmmwv_acc3/params"]
    n = 2000
    # zero controls from rest: every kinematic state stays exactly zero
    fin, _ = O.run18(flat, ("steps", np.zeros((n, 3, 1))), 1, n)
    kin = [60 + i for i in (0, 1, 2, 3, 4, 5, 6, 7)]
    assert np.all(fin[0, kin] == 0.0)
    # straight-line driving keeps the lateral / roll / yaw states exactly zero
    c = np.zeros((n, 3, 1))
    c[:, 1, 0] = 0.6
    fin, _ = O.run18(flat, ("steps", c), 1, n)
    v, psi, wz, phi, wx, y = fin[0, 63], fin[0, 64], fin[0, 65], fin[0, 66], fin[0, 67], fin[0, 61]
    assert fin[0, 62] > 1.0 and (v, psi, wz, phi, wx, y) == (0, 0, 0, 0, 0, 0)


def test_batch_independence(golden18):
    name = "hmmwv_knot_fleet"
    flat, ctrl, n, ovr = inputs(name, golden18)
    fin_all, _ = O.run18(flat, ctrl, n, 3000, ovr=ovr)
    _, ns, ds, np_, dp, knots = ctrl
    for i in (0, 5, 15):
        one = ("knots", ns, ds, np_, dp, knots[:, i:i + 1])
        fin1, _ = O.run18(flat, one, 1, 3000, ovr=ovr[:, i:i + 1])
        np.testing.assert_array_equal(fin1[0], fin_all[i])


def test_reference_model_is_ulp_sensitive_for_some_vehicles(golden18):
    """Documents why the 1e-9 parity bar is conditioned on ulp stability: the oracle (bit-identical
    to the reference) moves by far more than 1e-9 on a few vehicles of the knot_fleet workload when a
    single parameter moves by one ulp, while most vehicles stay within 1e-11."""
    from cases18 import ulp_spread
    from lfdm_amd import workloads as W

    flat = golden18["hmmwv_acc3/params"]
    n, nsteps = 512, 4000
    fl = W.knot_fleet(n, flat, seed=5)
    ctrl = ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    spread = ulp_spread(O, flat, ctrl, n, nsteps, fl["ovr"], 200)
    assert np.mean(spread <= 1e-11) > 0.9
    assert spread.max() > 1e-6


def test_flop_counting_matches_hand_counts():
    """The counting scalar behind every FLOP figure (oracle/orc_count.h) against operation counts made by
    hand from the reference source, so a dropped or double-counted operation would show.

    computeVehRHS (dof18.cpp:1019-1078), per statement (add, mul, div): mt (2, 1, 0), hrc (2, 2, 1),
    E1 (4, 2, 0), E2 (9, 8, 2), E3 (5, 8, 0), A1 (1, 2, 0), A2 (1, 2, 0), A3 (0, 1, 0), udot (7, 10, 1),
    denom (4, 11, 0), vdot / wxdot / wzdot (5, 12, 1) each: 50 add, 83 mul, 7 div.
    tmxy_combined (dof18.cpp:247-288): c = 2 fm / sm is (0, 1, 1) in every branch; then
      s < sm: p (1, 1, 1), sn (0, 0, 1), dn (2, 1, 0), f (0, 2, 1), fos (0, 0, 1)   -> 3 add, 5 mul, 5 div;
      s > ss: fos = fs / s (0, 0, 1)                                                 -> 0 add, 1 mul, 2 div;
      sm <= s <= sstar: a (0, 2, 3), sstar (3, 1, 1), f (3, 2, 0), fos (0, 0, 1)     -> 6 add, 6 mul, 6 div;
      sstar < s <= ss: a, sstar as above, b (2, 1, 1), f (3, 2, 0), fos (0, 0, 1)    -> 8 add, 7 mul, 7 div."""
    import lfdm_amd

    flat = lfdm_amd.params_from_json(18, f"{DATA}/18dof/HMMWV/vehicle.json", f"{DATA}/18dof/HMMWV/tmeasy.json")
    rec = np.zeros(75)
    rec[60 + 2], rec[60 + 3], rec[60 + 5], rec[60 + 6], rec[60 + 7] = 10.0, 0.2, 0.1, 0.01, 0.01
    assert O.count_veh_rhs18(flat, rec, [100, 200, 300, 400], [50, 60, 70, 80]) == (50, 83, 7, 0)
    # sm = 0.1, fm = 1000, ss = 0.5, fs = 800, df0 = 2e4: a = 5e4, sstar = 0.11
    args = (20000.0, 0.1, 1000.0, 0.5, 800.0)
    assert O.count_tmxy(0.05, *args) == (3, 5, 5, 0)
    assert O.count_tmxy(0.8, *args) == (0, 1, 2, 0)
    assert O.count_tmxy(0.105, *args) == (6, 6, 6, 0)
    assert O.count_tmxy(0.2, *args) == (8, 7, 7, 0)


def test_distinct_call_tallies():
    """The distinct-argument tallies count a libm call once per step per argument: the 24-DOF step calls
    sin / cos of pitch, roll, yaw and steering angle dozens of times, its distinct calls a few."""
    import lfdm_amd

    d = f"{DATA}/24dof/HMMWV"
    flat = lfdm_amd.params_from_json(24, f"{d}/vehicle.json", f"{d}/tmeasy.json", f"{d}/suspension.json")
    ctrl = np.tile(np.array([0.1, 0.5, 0.0])[None, :, None], (100, 1, 1))
    c = O.count_ops(24, flat, ("steps", ctrl), 100)
    assert c["sin"] >= 70 * 100 and c["cos"] >= 70 * 100
    assert c["sin_distinct"] <= 5 * 100 and c["cos_distinct"] <= 5 * 100
    for f in ("sin", "cos", "tan", "atan2", "hypot", "tanh"):
        assert c[f + "_distinct"] <= c[f]
