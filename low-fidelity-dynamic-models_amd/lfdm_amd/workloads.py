"""Seeded synthetic fleets for the parity and throughput workloads.

Inputs are generated once on the host with numpy's PCG64 so the GPU solver and the CPU
checkers consume identical doubles.  Steering given in radians is converted to the solver's
normalized input by dividing by the vehicle's maxSteer (linear steering).

Workloads (dt = 1e-3 s, HMMWV vehicle + TMeasy tire parameters, start at rest):
  ramp_sine  -- 18-DOF, 1,024 vehicles, 10 s: throttle ramps 0 -> U(0.3, 0.7) over 2 s, steering
                0.1 sin(2 pi 0.2 t + phi) rad with phi ~ U(0, 2 pi), no brake; per-step arrays.
  knot_fleet -- 18-DOF, 1,000,000 vehicles, 10 s: mass, yaw inertia and cornering stiffness
                jittered x U(0.9, 1.1); throttle piecewise constant U(0, 1) per 1 s; steering
                U(-0.3, 0.3) rad knots every 0.5 s, linearly interpolated; brake U(0, 0.5) on a
                seeded 10 % of the 1 s intervals; a knot program per vehicle.
  mixed_fleet -- 11/18/24-DOF, 2,000,000 vehicles split 40/40/20, per-vehicle horizons U(2, 20) s,
                parameters jittered x U(0.85, 1.15), knot programs; 24-DOF vehicles start at
                U(0, 20) m/s.
"""
import os

import numpy as np

from .solver import OVERRIDE_FIELDS

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DATA = os.path.join(ROOT, "tests", "golden", "data")


def data_path(*parts):
    return os.path.join(DATA, *parts)


HMMWV_VEH = data_path("18dof", "HMMWV", "vehicle.json")
HMMWV_TIRE = data_path("18dof", "HMMWV", "tmeasy.json")

# Index of fields in the flat 11-, 18- and 24-DOF parameter vectors (layouts in csrc/params.h);
# "gears" is the index of the gear count that opens the variable-length part.
FLAT18 = {"a": 0, "b": 1, "h": 2, "m": 3, "jz": 4, "jx": 5, "jxz": 6, "maxSteer": 17, "maxBrakeTorque": 19,
          "gears": 25}
FLAT11 = {"a": 0, "b": 1, "h": 2, "m": 3, "jz": 4, "maxSteer": 7, "maxBrakeTorque": 9, "gears": 15}
FLAT24 = {"a": 0, "b": 1, "h": 2, "m": 3, "jz": 4, "jx": 5, "jy": 6, "maxSteer": 13, "maxBrakeTorque": 15,
          "gears": 21}
FLAT = {11: FLAT11, 18: FLAT18, 24: FLAT24}


def time_grid(nsteps, dt, t0=0.0):
    """Step start times exactly as the solver accumulates them (t += dt)."""
    t = np.empty(nsteps)
    acc = t0
    for k in range(nsteps):
        t[k] = acc
        acc = acc + dt
    return t


def tire_offsets(flat, tire_nr=False, dof=18):
    """Index of dfy0Pn / dfy0P2n in a flat parameter vector."""
    i = FLAT[dof]["gears"]
    ng = int(flat[i])
    i += 1 + ng
    for _ in range(6):
        n = int(flat[i])
        i += 1 + 2 * n
    base = i
    if not tire_nr:
        return base + 25, base + 26
    return base + 23, base + 24


def default_overrides(flat, n, tire_nr=False, dof=18):
    """(7, n) per-vehicle parameter array holding the group's own values (jx is 0 for the 11-DOF
    model, which has no roll)."""
    F = FLAT[dof]
    i0, i1 = tire_offsets(flat, tire_nr, dof)
    vals = [flat[F["m"]], flat[F["jz"]], flat[F["jx"]] if "jx" in F else 0.0, flat[i0], flat[i1], 1.0,
            flat[F["maxBrakeTorque"]]]
    return np.repeat(np.array(vals)[:, None], n, axis=1)


def ramp_sine(n, nsteps=10000, dt=1e-3, seed=1, max_steer=0.582178801):
    """Per-step controls (nsteps, 3, n) of the ramp_sine workload."""
    rng = np.random.Generator(np.random.PCG64(seed))
    c = rng.uniform(0.3, 0.7, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    t = time_grid(nsteps, dt)
    ctrl = np.empty((nsteps, 3, n))
    ctrl[:, 0, :] = (0.1 * np.sin(2 * np.pi * 0.2 * t[:, None] + phi[None, :])) / max_steer
    ctrl[:, 1, :] = c[None, :] * np.minimum(t[:, None] / 2.0, 1.0)
    ctrl[:, 2, :] = 0.0
    return ctrl


def knot_fleet(n, flat, seed=2, horizon=10.0, max_steer=None, jitter=0.1, dof=18, tire_nr=False):
    """Knot program and per-vehicle parameters of the knot_fleet workload.

    Returns dict(ns, ds, np, dp, knots (ns + 2 np, n), ovr (7, n))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    max_steer = flat[FLAT[dof]["maxSteer"]] if max_steer is None else max_steer
    ds, dp = 0.5, 1.0
    ns = int(round(horizon / ds)) + 1
    np_ = int(round(horizon / dp))
    steer = rng.uniform(-0.3, 0.3, (ns, n)) / max_steer
    thr = rng.uniform(0.0, 1.0, (np_, n))
    brk = np.where(rng.random((np_, n)) < 0.1, rng.uniform(0.0, 0.5, (np_, n)), 0.0)
    knots = np.concatenate([steer, thr, brk], axis=0)
    ovr = default_overrides(flat, n, tire_nr, dof)
    sm = rng.uniform(1 - jitter, 1 + jitter, n)
    sj = rng.uniform(1 - jitter, 1 + jitter, n)
    sc = rng.uniform(1 - jitter, 1 + jitter, n)
    ovr[0] = ovr[0] * sm
    ovr[1] = ovr[1] * sj
    ovr[3] = ovr[3] * sc
    ovr[4] = ovr[4] * sc
    return dict(ns=ns, ds=ds, np=np_, dp=dp, knots=np.ascontiguousarray(knots), ovr=np.ascontiguousarray(ovr))


def initial_speed_records(n, flat, dof=24, seed=7, umax=20.0, tire_nr=False):
    """Initial records (n, rec) of a fleet that starts moving: forward speed u0 ~ U(0, umax) per
    vehicle, every wheel rolling at u0 / r0, first gear, crank at the wheel speed over the first
    gear ratio (the transmission's output-to-input relation, omega_out = omega_wheel / ratio).
    Everything else is zero; for the 24-DOF model Initialize then adds the static tire and strut
    state (initializeTireSus)."""
    from .solver import record_fields

    rng = np.random.Generator(np.random.PCG64(seed))
    u0 = rng.uniform(0.0, umax, n)
    F = FLAT[dof]
    gr0 = flat[F["gears"] + 1]
    r0 = flat[tire_offsets(flat, tire_nr, dof)[0] - 25 + 3]  # tire block: jw rr mu r0 ...
    fields = record_fields(dof, 1 if tire_nr else 0)
    rec = np.zeros((n, len(fields)))
    rec[:, fields.index("u")] = u0
    om = u0 / r0
    for i, f in enumerate(fields):
        if f.endswith(".omega"):
            rec[:, i] = om
    rec[:, fields.index("crankOmega")] = om / gr0
    return rec


def mixed_fleet(n_total, split=(0.4, 0.4, 0.2), tmin=2.0, tmax=20.0, jitter=0.15, seed=11, dt=1e-3):
    """Per-model inputs of the mixed fleet (batching form), one model at a time: n_total vehicles split
    over the 11-, 18- and 24-DOF HMMWV models, mass / yaw inertia / cornering stiffness jittered
    x U(1 - jitter, 1 + jitter), knot-program controls over tmax, per-vehicle horizons
    T_i ~ U(tmin, tmax) s (steps_i = round(T_i / dt)); 11/18-DOF vehicles start at rest, 24-DOF ones at
    U(0, 20) m/s.  Yields dict(dof, n, flat, fl (knot_fleet), recs (24-DOF initial records or None),
    horizon (int32 steps))."""
    from . import params_from_json

    rng = np.random.Generator(np.random.PCG64(seed))
    for k, (dof, frac) in enumerate(zip((11, 18, 24), split)):
        n = int(round(n_total * frac))
        d = os.path.join(DATA, f"{dof}dof", "HMMWV")
        flat = params_from_json(dof, f"{d}/vehicle.json", f"{d}/tmeasy.json",
                                f"{d}/suspension.json" if dof == 24 else None)
        fl = knot_fleet(n, flat, seed=seed + k, horizon=tmax, jitter=jitter, dof=dof)
        recs = initial_speed_records(n, flat, dof, seed=seed + k) if dof == 24 else None
        hz = np.rint(rng.uniform(tmin, tmax, n) / dt).astype(np.int32)
        yield dict(dof=dof, n=n, flat=flat, fl=fl, recs=recs, horizon=hz)


def knot_controls(t, ns, ds, np_, dp, knots):
    """Per-step controls (len(t), 3, n) of a knot program, evaluated with the solver's formula:
    steering lerped on the ds grid (segment index floor(t / ds) clamped to [0, ns - 2], weight
    clamped to <= 1), throttle and braking taken from interval floor(t / dp) clamped to [0, np - 1]."""
    t = np.asarray(t, dtype=np.float64)
    ks = np.clip(np.floor(t / ds).astype(np.int64), 0, ns - 2)
    kp = np.clip(np.floor(t / dp).astype(np.int64), 0, np_ - 1)
    tk = ks * ds
    tbar = np.minimum((t - tk) / ds, 1.0)
    s0 = knots[ks]
    s1 = knots[ks + 1]
    out = np.empty((t.size, 3, knots.shape[1]))
    out[:, 0, :] = s0 + tbar[:, None] * (s1 - s0)
    out[:, 1, :] = knots[ns + kp]
    out[:, 2, :] = knots[ns + np_ + kp]
    return out


assert len(OVERRIDE_FIELDS) == 7
