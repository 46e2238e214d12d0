"""The solver's step code compiled for the host (tests/cpp/host_step.cpp) in both arithmetic forms,
against the oracle.

On the host every quotient is a division and every transcendental glibc's, so the strict form
(LFDM_ARITH_STRICT) must reproduce the oracle, and with it the reference CPU solver, bit for bit: that
pins the strict step's formulas (the reference's quotients, atan2 / tan slip angle, hypot, (dx num) /
den, and which sine / cosine pairs the reference's compiled functions take from one sincos) without a
GPU.  The fast form's reformulations (tangent-difference slip angle, sqrt for hypot, reciprocals of
shared denominators) change only rounding, so the fast form is held to the fleet parity bar of
fleetcheck.py: 1e-9 for vehicles whose trajectory does not respond to a one-ulp parameter change, the
model's own one-ulp response (times fleetcheck.SPREAD_MULT) for the others.  These fleets include
ill-conditioned vehicles (the Sedan's 24-DOF TMeasyNr vehicles started at speed pitch and roll hard).
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import DATA, ROOT

import fleetcheck as F
import lfdm_amd
import lfdm_oracle as O
from lfdm_amd import workloads as W

LIB = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "libhost_step.so")

CASES = [  # dof, vehicle, tire, TMeasyNr?, suspension
    (18, "HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, None),
    (18, "Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, None),
    (18, "ART/vehicle.json", "ART/tmeasy.json", False, None),
    (11, "HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, None),
    (11, "Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, None),
    (24, "HMMWV/vehicle.json", "HMMWV/tmeasy.json", False, "HMMWV/suspension.json"),
    (24, "Sedan/vehicle.json", "Sedan/tmeasyNr.json", True, "Sedan/suspension.json"),
]


def host_lib():
    assert os.path.exists(LIB), "run make -C low-fidelity-dynamic-models_amd hoststep"
    L = ctypes.CDLL(LIB)
    i, p, d = ctypes.c_int, ctypes.c_void_p, ctypes.c_double
    L.host_step_run.argtypes = [i, i, i, p, i, p, p, i, i, i, p, d, p, ctypes.c_char_p, i]
    L.host_step_run.restype = i
    return L


def host_run(L, dof, nr, strict, flat, init, ctrl, ovr):
    n = ctrl.shape[2]
    fin = np.zeros_like(init)
    err = ctypes.create_string_buffer(256)
    rc = L.host_step_run(dof, int(nr), int(strict), flat.ctypes.data, flat.size, ovr.ctypes.data, init.ctypes.data, n,
                         init.shape[1], ctrl.shape[0], ctrl.ctypes.data, 1e-3, fin.ctypes.data, err, 256)
    assert rc == 0, err.value
    return fin


def host_sampled(L, dof, nr, strict, flat, init, ctrl, ovr, every):
    """Records after every `every` steps (chunked runs continue from the records, as launches do)."""
    out, cur = [], init
    for k in range(ctrl.shape[0] // every):
        cur = host_run(L, dof, nr, strict, flat, cur, np.ascontiguousarray(ctrl[k * every:(k + 1) * every]), ovr)
        out.append(cur)
    return np.stack(out)


@pytest.mark.parametrize("dof,veh,tire,nr,sus", CASES, ids=[f"{c[0]}-{c[1].split('/')[0]}-{c[2]}" for c in CASES])
def test_host_step_forms_vs_oracle(dof, veh, tire, nr, sus):
    L = host_lib()
    d = os.path.join(DATA, f"{dof}dof")
    flat = lfdm_amd.params_from_json(dof, f"{d}/{veh}", f"{d}/{tire}", f"{d}/{sus}" if sus else None, int(nr))
    n, nsteps, every = 48, 10000, 500
    # steering 0.1 sin(...) rad through a linear map of maxSteer 0.3
    ctrl = np.ascontiguousarray(W.ramp_sine(n, nsteps=nsteps, seed=3, max_steer=0.3))
    ovr = W.knot_fleet(n, flat, seed=4, dof=dof, tire_nr=nr)["ovr"]
    recs = W.initial_speed_records(n, flat, dof, seed=6, tire_nr=nr) if dof == 24 else None
    case = dict(dof=dof, nr=nr, n=n, flat=flat, fl=dict(ovr=ovr), ctrl=("steps", ctrl), recs=recs)
    init = F.oracle_init(O, case, ovr)
    init = np.zeros((n, O.record_len(dof, nr))) if init is None else init
    R = F.reference(O, case, nsteps, every)
    strict = host_sampled(L, dof, nr, True, flat, init, ctrl, ovr, every)
    np.testing.assert_array_equal(strict, R.traj)
    fast = host_sampled(L, dof, nr, False, flat, init, ctrl, ovr, every)
    v = F.verdict(fast, R)
    print(F.summary_line(f"{dof}-DOF {veh} {tire}, fast form", v))
    assert v["ok"], v["failing"]
    assert not np.array_equal(fast, R.traj)


@pytest.mark.parametrize("dof", [18, 11, 24])
def test_host_step_forms_at_rest_golden(dof):
    """The reference's step_api case (standing still for 0.5 s, then steering 0.1 and throttle 0.5) through the
    host step code, against the reference's trajectory (golden fixture), both forms within 1e-9 and the strict
    form bit for bit.  Standing still, every force of the reference model is exactly 0 (the powertrain's loss
    map passes through the origin, the tire deflections stay 0), and the rolling-resistance moment flips with
    the sign of the wheel speed: a form that leaves a rounding residue where the reference has an exact 0
    (a reciprocal segment width in a map lookup does) starts that flip and departs at O(1)."""
    import cases18
    import cases1124

    L = host_lib()
    g = np.load(os.path.join(ROOT, "tests", "golden", f"d{dof}_golden.npz"), allow_pickle=False)
    name = "hmmwv_step_api"
    if dof == 18:
        flat, ctrl, n, _ = cases18.inputs(name, g)
        init = np.zeros((1, O.record_len(18, False)))
    else:
        flat, ctrl, n, _, init = cases1124.inputs(dof, name, g)
        init = np.ascontiguousarray(init.reshape(1, -1))
    nsteps, every = (int(v) for v in g[f"{name}/spec"][:2])
    c = np.ascontiguousarray(ctrl[1])
    ovr = W.default_overrides(flat, 1, dof=dof)
    ref = g[f"{name}/traj"]
    for strict in (True, False):
        tr = host_sampled(L, dof, False, strict, flat, init, c, ovr, every)
        err = cases18.normwise_rel(tr, ref.reshape(tr.shape))
        print(f"{dof}-DOF {name}, {'strict' if strict else 'fast'} form: max normwise rel {err.max():.3e}")
        if strict:
            np.testing.assert_array_equal(tr, ref.reshape(tr.shape))
        assert err.max() <= F.TOL


def test_fast_sines_and_cosines_on_the_host():
    """sincos_fast (Cody-Waite reduction, Taylor polynomials) and sin_2pi_unit (exact reduction of a turn
    fraction) as the host build evaluates them, against x86 long double sinl / cosl rounded to double:
    within two ulps over steering-sized, multi-turn and long-run (1e5 rad) angles and over [0, 1] turns,
    exact at the quarter turns; NaN stays NaN.  (tests/test_gpu_arith.py checks the device build.)"""
    L = host_lib()
    L.host_fast_trig.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    rng = np.random.default_rng(9)
    n = 30000
    a = np.concatenate([rng.uniform(-0.6, 0.6, n // 3), rng.uniform(-20, 20, n // 3), rng.uniform(-1e5, 1e5, n // 3)])
    v = np.concatenate([rng.uniform(0, 1, len(a) - 5), [0.0, 0.25, 0.5, 0.75, 1.0]])
    out = np.empty(3 * len(a))
    L.host_fast_trig(a.ctypes.data, v.ctypes.data, len(a), out.ctypes.data)
    s, c, s2 = out[:len(a)], out[len(a):2 * len(a)], out[2 * len(a):]
    ld = np.longdouble
    ref_s = np.sin(a.astype(ld)).astype(np.float64)
    ref_c = np.cos(a.astype(ld)).astype(np.float64)
    # sin(2 pi v) = sin(2 pi t) for the exactly reduced t, whose long double product with 2 pi does not cancel
    r = v - np.rint(v)
    t = np.where(np.abs(r) > 0.25, np.where(r < 0, -0.5, 0.5) - r, r)
    two_pi = ld(2) * ld(np.pi) + ld(2) * (ld(3.141592653589793238462643383279502884) - ld(np.pi))
    ref_s2 = np.sin(two_pi * t.astype(ld)).astype(np.float64)

    def ulps(x, ref):
        return np.abs(x - ref) / np.spacing(np.abs(ref))

    assert ulps(s, ref_s).max() <= 2 and ulps(c, ref_c).max() <= 2, (ulps(s, ref_s).max(), ulps(c, ref_c).max())
    nz = ref_s2 != 0
    assert ulps(s2[nz], ref_s2[nz]).max() <= 2
    np.testing.assert_array_equal(s2[-5:], [0.0, 1.0, 0.0, -1.0, 0.0])
    nan = np.array([np.nan])
    o = np.empty(3)
    L.host_fast_trig(nan.ctypes.data, nan.ctypes.data, 1, o.ctypes.data)
    assert np.isnan(o).all()


def test_fast_sines_and_cosines_of_huge_angles_on_the_host():
    """sincos_fast beyond its Cody-Waite range (|a| >= 2^20 pi/2) reduces by Payne-Hanek with a 2/pi
    table: against mpmath at 2,000 bits on log-uniform angles up to 2^1024 and on the double closest to
    a multiple of pi/2 (6381956970095103 2^797, where cos is 4.7e-19), within one ulp; NaN for inf, NaN;
    and the table's 37 words are the bits of 2/pi (Machin's formula in integer arithmetic)."""
    mpmath = pytest.importorskip("mpmath")
    L = host_lib()
    L.host_fast_trig.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.host_two_over_pi_word.restype = ctypes.c_uint
    L.host_two_over_pi_word.argtypes = [ctypes.c_int]

    def atan_inv(x, one):  # atan(1/x) * one
        s, t, k, sign = 0, one // x, 1, 1
        while t:
            s += sign * (t // k)
            t //= x * x
            k += 2
            sign = -sign
        return s
    prec = 1400
    one = 1 << prec
    pi = 16 * atan_inv(5, one) - 4 * atan_inv(239, one)
    bits = (2 << (prec + 37 * 32)) // pi  # floor(2/pi 2^1184), exact but for the last few bits
    words = [(bits >> (32 * (36 - k))) & 0xffffffff for k in range(37)]
    assert [L.host_two_over_pi_word(k) for k in range(39)] == [0, 0] + words

    rng = np.random.default_rng(5)
    n = 3000
    a = np.exp2(rng.uniform(20.66, 1023.99, n)) * rng.choice([-1.0, 1.0], n)
    a = np.concatenate([a, [6381956970095103 * 2.0 ** 797, -6381956970095103 * 2.0 ** 797, 2.0 ** 1023,
                            np.finfo(np.float64).max, float.fromhex("0x1.921fb54442d18p+20")]])
    out = np.empty(3 * len(a))
    L.host_fast_trig(a.ctypes.data, a.ctypes.data, len(a), out.ctypes.data)
    s, c = out[:len(a)], out[len(a):2 * len(a)]
    mpmath.mp.prec = 2000
    rs = np.array([float(mpmath.sin(mpmath.mpf(float(x)))) for x in a])
    rc = np.array([float(mpmath.cos(mpmath.mpf(float(x)))) for x in a])
    us = np.abs(s - rs) / np.spacing(np.abs(rs))
    uc = np.abs(c - rc) / np.spacing(np.abs(rc))
    assert us.max() <= 1 and uc.max() <= 1, (us.max(), uc.max(), a[us.argmax()], a[uc.argmax()])
    nf = np.array([np.inf, -np.inf, np.nan])
    o = np.empty(9)
    L.host_fast_trig(nf.ctypes.data, nf.ctypes.data, 3, o.ctypes.data)
    assert np.isnan(o[:6]).all()
