"""Fleet-scale parity of the HIP solver against the oracle: the seeded fleets of the parity tests, their
GPU runs in both arithmetic forms, the oracle's reference trajectories with the model's own one-ulp
sensitivity, and the per-vehicle verdict.

Why a sensitivity term: the reference model has discrete events (gear shifts, sgn(omega) of the brake
torque at a locked wheel, clamps, the tire's off-ground reset), so for some vehicles a one-ulp change
anywhere grows to a visible trajectory change.  The device libm and glibc round some arguments
differently, so such vehicles cannot be held to 1e-9.  The oracle measures how sensitive each vehicle
is: its trajectory with one of its parameters moved by one ulp (m, Izz, Ixx, dfy0Pn, torque scale and
brake torque, each up and down) against the unperturbed one.  The spread is the largest of these
normwise changes.  The bar, per vehicle:
  - spread <= STABLE (1e-11): the GPU trajectory is within TOL = 1e-9 of the oracle;
  - otherwise within max(TOL, SPREAD_MULT x spread): the GPU's difference is of the size of the
    model's own response to one ulp.  SPREAD_MULT = 25: on MI355X, over 10 s, the largest ratio of GPU
    error to spread was 12 (over the 18-, 11- and 24-DOF TMeasy and TMeasyNr fleets, the mixed fleet and
    both arithmetic forms; 99 % of the ill-conditioned vehicles stay below 7);
  - non-finite values (the reference model blows up for a few Sedan 24-DOF TMeasyNr vehicles) occur in
    the GPU trajectory exactly for the vehicles whose oracle trajectory has them, wherever that is
    itself ulp-stable (the base run and every one-ulp run agree on it); a vehicle whose blow-up comes
    and goes with one ulp may go either way.  Errors cover the samples before the first non-finite
    one of either trajectory.
Errors are normwise: max over samples of |a - b| / max over samples of |b|, per vehicle and record
field (0 / 0 counts as 0), maximum over fields.
"""
import numpy as np

from lfdm_amd import ARITH_FAST, ARITH_STRICT, SolverHalfImplicitGPU, TMeasy, TMeasyNr
from lfdm_amd import workloads as W

TOL = 1e-9
STABLE = 1e-11
SPREAD_MULT = 25.0
# Per-vehicle parameters moved by one ulp for the sensitivity (indices into the (7, n) array: m, jz,
# jx, dfy0Pn, torque scale, maxBrakeTorque); the 11-DOF model has no roll inertia.
ULP_FIELDS = {11: (0, 1, 3, 5, 6), 18: (0, 1, 2, 3, 5, 6), 24: (0, 1, 2, 3, 5, 6)}


def build_fleet(dof, nr, n, flat, seed, init_seed=7):
    """The knot_fleet generator (per-vehicle m, Izz, cornering stiffness x U(0.9, 1.1); throttle
    piecewise constant U(0, 1) per 1 s; steering U(-0.3, 0.3) rad knots per 0.5 s; brake U(0, 0.5) on
    10 % of the 1 s intervals).  24-DOF vehicles start at U(0, 20) m/s (W.initial_speed_records).
    Returns a dict with the oracle's inputs."""
    fl = W.knot_fleet(n, flat, seed=seed, dof=dof, tire_nr=nr)
    ctrl = ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    recs = W.initial_speed_records(n, flat, dof, seed=init_seed, tire_nr=nr) if dof == 24 else None
    return dict(dof=dof, nr=nr, n=n, flat=flat, fl=fl, ctrl=ctrl, recs=recs)


def oracle_init(O, case, ovr):
    """Initial records of the oracle run for per-vehicle parameters ovr (24-DOF: initializeTireSus,
    which depends on the vehicle's mass)."""
    if case["dof"] != 24:
        return None
    return O.init24(case["flat"], case["recs"], case["nr"], ovr)


def gpu_solver(case, arith=ARITH_FAST):
    n, dof, nr = case["n"], case["dof"], case["nr"]
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(case["flat"], n, None, TMeasyNr if nr else TMeasy)
    s.set_overrides(case["fl"]["ovr"])
    if dof == 24:
        s.initialize_batch(case["recs"])
    else:
        s.Initialize(s.zero_record(), n)
    fl = case["fl"]
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    s.set_arithmetic(arith)
    return s


def gpu_run(case, nsteps, every, arith=ARITH_FAST):
    """Records of every vehicle after every `every` steps, (nsteps // every, n, rec)."""
    s = gpu_solver(case, arith)
    out = []
    for _ in range(nsteps // every):
        s.advance(every)
        out.append(s.get_states())
    s.close()
    return np.stack(out)


def oracle_run(O, case, nsteps, every, ovr=None):
    ovr = case["fl"]["ovr"] if ovr is None else ovr
    _, tr = O.run(case["dof"], case["flat"], case["ctrl"], case["n"], nsteps, tire_nr=case["nr"], ovr=ovr,
                  init=oracle_init(O, case, ovr), every=every)
    return tr


def finite_prefix(traj):
    """Per vehicle, the number of leading samples whose records are all finite."""
    bad = ~np.isfinite(traj).all(axis=2)  # (samples, n)
    first = np.where(bad.any(axis=0), bad.argmax(axis=0), traj.shape[0])
    return first


def normwise(a, b, upto):
    """Per vehicle: max over fields of max_t<upto |a - b| / max_t<upto |b| (0 / 0 = 0)."""
    ns = a.shape[0]
    keep = (np.arange(ns)[:, None] < upto[None, :])[:, :, None]
    d = np.where(keep, np.abs(a - b), 0.0)
    m = np.where(keep, np.abs(b), 0.0)
    d = np.nan_to_num(d, nan=np.inf, posinf=np.inf).max(axis=0)
    m = m.max(axis=0)
    r = np.where(m > 0, d / np.where(m > 0, m, 1), d)
    return r.max(axis=-1)


class Reference:
    """The oracle's trajectory `traj`, each vehicle's finite prefix `upto`, its one-ulp `spread` over the
    common finite prefix, and `nf_flip`: vehicles whose finiteness changes under a one-ulp change."""

    def __init__(self, traj, upto, spread, nf_flip):
        self.traj, self.upto, self.spread, self.nf_flip = traj, upto, spread, nf_flip


def reference(O, case, nsteps, every):
    ref = oracle_run(O, case, nsteps, every)
    upto = finite_prefix(ref)
    nf = upto < ref.shape[0]
    spread = np.zeros(case["n"])
    nf_flip = np.zeros(case["n"], dtype=bool)
    for f in ULP_FIELDS[case["dof"]]:
        for direction in (np.inf, -np.inf):
            o2 = case["fl"]["ovr"].copy()
            o2[f] = np.nextafter(o2[f], direction)
            tr = oracle_run(O, case, nsteps, every, o2)
            up2 = finite_prefix(tr)
            nf_flip |= (up2 < tr.shape[0]) != nf
            spread = np.maximum(spread, normwise(tr, ref, np.minimum(upto, up2)))
    return Reference(ref, upto, spread, nf_flip)


def _assess(err, spread, gpu_nf, ref_nf, nf_flip, mult):
    stable = spread <= STABLE
    with np.errstate(over="ignore", invalid="ignore"):
        bound = np.where(stable, TOL, np.maximum(TOL, mult * spread))
        ratio = np.where(~stable & (spread > 0), err / np.where(spread > 0, spread, 1), 0.0)
    nf_ok = (gpu_nf == ref_nf) | nf_flip
    ok_v = (err <= bound) & nf_ok
    finite_ratio = ratio[~stable & np.isfinite(ratio)]
    return stable, ok_v, nf_ok, finite_ratio


def verdict(gpu, R, mult=None):
    """Per-vehicle errors and the bar of this module against Reference R; returns a summary dict (ok:
    every vehicle meets the bar)."""
    mult = SPREAD_MULT if mult is None else mult
    ref, spread = R.traj, R.spread
    n = ref.shape[1]
    gup = finite_prefix(gpu)
    gpu_nonfinite = gup < gpu.shape[0]
    ref_nonfinite = R.upto < ref.shape[0]
    err = normwise(gpu, ref, np.minimum(R.upto, gup))
    stable, ok_v, nf_ok, fr = _assess(err, spread, gpu_nonfinite, ref_nonfinite, R.nf_flip, mult)
    fin = np.isfinite(ref).all(axis=2) & np.isfinite(gpu).all(axis=2)
    return dict(
        n=n,
        ok=bool(ok_v.all()),
        nonfinite_match=bool(nf_ok.all()),
        nonfinite=int(ref_nonfinite.sum()),
        nonfinite_gpu=int(gpu_nonfinite.sum()),
        nonfinite_flip=int(R.nf_flip.sum()),
        stable=int(stable.sum()),
        max_err_stable=float(err[stable].max()) if stable.any() else 0.0,
        outside_tol=int((err > TOL).sum()),
        failing=np.flatnonzero(~ok_v).tolist(),
        max_ratio=float(fr.max()) if fr.size else 0.0,
        ratio_q=[float(q) for q in np.quantile(fr, [0.5, 0.9, 0.99, 1.0])] if fr.size else [],
        bitwise=float(np.mean(gpu[fin] == ref[fin])),
    )


def summary_line(tag, v):
    return (f"{tag}: {v['n']} vehicles, {v['stable']} ulp-stable (max rel {v['max_err_stable']:.2e}); "
            f"{v['outside_tol']} outside 1e-9; err/spread of the others median/p90/p99/max "
            f"{', '.join(f'{q:.2g}' for q in v['ratio_q'])}; non-finite: oracle {v['nonfinite']}, GPU "
            f"{v['nonfinite_gpu']}, ulp-unstable {v['nonfinite_flip']} (rule met {v['nonfinite_match']}); bitwise "
            f"fraction {v['bitwise']:.4f}; failing {v['failing'][:10]}")


ARITHS = {"fast": ARITH_FAST, "strict": ARITH_STRICT}


# ---- the mixed fleet (batching form): final records at each vehicle's own horizon ----

def mixed_inputs(n_total, seed=11):
    return list(W.mixed_fleet(n_total, seed=seed))


def mixed_gpu(models, arith=ARITH_FAST, trial_steps=200):
    """Runs the mixed fleet as one graph launch per Fleet.run (autotuned); returns the final records per
    model and the tuned configuration."""
    from lfdm_amd import Fleet

    buckets = []
    for m in models:
        s = SolverHalfImplicitGPU(m["dof"], m["n"])
        s.construct_flat(m["flat"], m["n"], None, TMeasy)
        s.set_overrides(m["fl"]["ovr"])
        if m["dof"] == 24:
            s.initialize_batch(m["recs"])
        else:
            s.Initialize(s.zero_record(), m["n"])
        fl = m["fl"]
        s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
        s.set_horizons(m["horizon"])
        s.set_arithmetic(arith)
        buckets.append(s)
    T = int(max(m["horizon"].max() for m in models))
    f = Fleet(buckets)
    init = [s.get_states() for s in buckets]
    cfg = f.autotune(T, trial_steps=trial_steps)
    for s, i0 in zip(buckets, init):
        np.testing.assert_array_equal(s.get_states(), i0)  # the autotuner restores the records
    f.run(T)
    f.synchronize()
    fins = [s.get_states() for s in buckets]
    f.close()
    for s in buckets:
        s.close()
    return fins, cfg


def final_reference(O, dof, flat, ctrl, n, nsteps, ovr, init_fn=lambda o: None, horizon=None, every=500,
                    tire_nr=False):
    """Oracle final records (Reference.traj) of n vehicles after nsteps (or each vehicle's horizon),
    normwise denominators `den` from the trajectory sampled every `every` steps up to the end, the one-ulp
    spread of the final records and the finiteness flips.  init_fn(ovr) gives the initial records."""

    def run(o):
        return O.run(dof, flat, ctrl, n, nsteps, tire_nr=tire_nr, ovr=o, init=init_fn(o), every=every,
                     horizon=horizon)

    fin, tr = run(ovr)
    with np.errstate(invalid="ignore"):
        den = np.maximum(np.nan_to_num(np.abs(tr), nan=0.0).max(axis=0), np.nan_to_num(np.abs(fin), nan=0.0))
    nf = ~np.isfinite(fin).all(axis=-1)
    spread = np.zeros(n)
    nf_flip = np.zeros(n, dtype=bool)
    for f in ULP_FIELDS[dof]:
        for direction in (np.inf, -np.inf):
            o2 = ovr.copy()
            o2[f] = np.nextafter(o2[f], direction)
            fin2, _ = run(o2)
            nf2 = ~np.isfinite(fin2).all(axis=-1)
            nf_flip |= nf2 != nf
            both = ~(nf | nf2)
            spread = np.maximum(spread, np.where(both, final_err(fin2, fin, den), np.where(nf2 != nf, np.inf, 0.0)))
    R = Reference(fin, None, spread, nf_flip)
    R.den = den
    return R


def mixed_reference(O, m, every=500):
    """final_reference for one model of the mixed fleet, to each vehicle's horizon."""
    dof, n, flat, fl = m["dof"], m["n"], m["flat"], m["fl"]
    ctrl = ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    T = int(m["horizon"].max())
    init_fn = (lambda o: O.init24(flat, m["recs"], False, o)) if dof == 24 else (lambda o: None)
    return final_reference(O, dof, flat, ctrl, n, T, fl["ovr"], init_fn, m["horizon"], every)


def final_err(a, b, den):
    """Per vehicle: max over fields of |a - b| / den (0 / 0 = 0; non-finite differences are infinite)."""
    with np.errstate(invalid="ignore"):
        d = np.nan_to_num(np.abs(a - b), nan=np.inf, posinf=np.inf)
        r = np.where(den > 0, d / np.where(den > 0, den, 1), d)
    return r.max(axis=-1)


def final_verdict(gpu, R, mult=None):
    """verdict() for final records only (the mixed fleet's graph runs to each vehicle's horizon)."""
    mult = SPREAD_MULT if mult is None else mult
    ref, spread = R.traj, R.spread
    ref_nf = ~np.isfinite(ref).all(axis=-1)
    gpu_nf = ~np.isfinite(gpu).all(axis=-1)
    both = ~(ref_nf | gpu_nf)
    err = np.where(both, final_err(np.where(both[:, None], gpu, 0), np.where(both[:, None], ref, 0), R.den), 0.0)
    stable, ok_v, nf_ok, fr = _assess(err, spread, gpu_nf, ref_nf, R.nf_flip, mult)
    return dict(
        n=ref.shape[0],
        ok=bool(ok_v.all()),
        nonfinite_match=bool(nf_ok.all()),
        nonfinite=int(ref_nf.sum()),
        nonfinite_gpu=int(gpu_nf.sum()),
        nonfinite_flip=int(R.nf_flip.sum()),
        stable=int(stable.sum()),
        max_err_stable=float(err[stable].max()) if stable.any() else 0.0,
        outside_tol=int((err > TOL).sum()),
        failing=np.flatnonzero(~ok_v).tolist(),
        max_ratio=float(fr.max()) if fr.size else 0.0,
        ratio_q=[float(q) for q in np.quantile(fr, [0.5, 0.9, 0.99, 1.0])] if fr.size else [],
        bitwise=float(np.mean(gpu[both] == ref[both])),
    )
