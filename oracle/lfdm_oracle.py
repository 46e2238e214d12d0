"""ORACLE — test infrastructure only.

Python binding of oracle/liblfdm_oracle.so, the CPU restatement of the reference half-implicit
solvers (wheeled_vehicle_models/{11,18,24}dof/dof{11,18,24}.cpp and dof*_halfImplicit.cpp).  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and only as the
checker or the CPU baseline, never as the thing measured.

Pinning (see tests/test_oracle.py): the 18-DOF restatement is bit-identical to the reference
solver compiled from its own sources (oracle/_ref) on the golden trajectories in tests/golden/
(HMMWV TMeasy, Sedan TMeasyNr, ART TMeasy with throttleMod, audi RWD; driver files and random
per-step controls), and checked against the reference's own known-answer values.  The 11- and
24-DOF restatements are pinned the same way (tests/test_oracle1124.py, tests/golden/d11_golden.npz,
d24_golden.npz).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblfdm_oracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        p, d, i = ctypes.c_void_p, ctypes.c_double, ctypes.c_int
        L.orc18_record_len.argtypes = [i]
        L.orc18_record_len.restype = i
        for m in ("11", "18", "24"):
            getattr(L, f"orc{m}_record_len").argtypes = [i]
            getattr(L, f"orc{m}_record_len").restype = i
            getattr(L, f"orc{m}_run").argtypes = [i, p, p, p, i, d, d, i, i, p, p, p, i, p, p, i]
            getattr(L, f"orc{m}_run").restype = i
        L.orc_run_h.argtypes = [i, i, p, p, p, i, d, d, i, p, i, p, p, p, i, p, p, i]
        L.orc_run_h.restype = i
        L.orc18_loss_tangents.argtypes = [i, p, p, p, i, i, d, p, p, p, p, p, p, p, i]
        L.orc18_loss_tangents.restype = i
        L.orc_dual_op.argtypes = [i, d, d, p, p]
        L.orc_dual_op.restype = d
        L.orc24_init.argtypes = [i, p, p, i, p]
        L.orc24_init.restype = i
        L.orc18_count_ops.argtypes = [i, p, p, d, d, i, i, p, p, p, p]
        L.orc18_count_ops.restype = i
        L.orc_count_ops.argtypes = [i, i, p, p, d, d, i, i, p, p, p, p]
        L.orc_count_ops.restype = i
        L.orc_count_veh_rhs18.argtypes = [p, p, p, p, p]
        L.orc_count_veh_rhs18.restype = None
        L.orc_count_tmxy.argtypes = [d, d, d, d, d, d, p]
        L.orc_count_tmxy.restype = None
        L.orc_tmxy.argtypes = [d, d, d, d, d, d, p, p]
        L.orc_tmxy.restype = None
        for fn in ("orc_interp_l", "orc_interp_q"):
            getattr(L, fn).argtypes = [d, d, d, d]
            getattr(L, fn).restype = d
        L.orc_sine_step.argtypes = [d, d, d, d, d]
        L.orc_sine_step.restype = d
        L.orc_map_y.argtypes = [p, i, d]
        L.orc_map_y.restype = d
        L.orc_diff_split.argtypes = [d, d, d, p, p]
        L.orc_diff_split.restype = None
        L.orc18_loads.argtypes = [i, p, p, p]
        L.orc18_loads.restype = None
        L.orc18_veh_rhs.argtypes = [i, p, p, p, p, p]
        L.orc18_veh_rhs.restype = None
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def _ctrl(mode_args):
    """(mode, ctrl array, imeta, dmeta) from a control description."""
    kind = mode_args[0]
    if kind == "steps":
        return 0, np.ascontiguousarray(mode_args[1], dtype=np.float64), None, None
    if kind == "table":
        rows = np.ascontiguousarray(mode_args[1], dtype=np.float64)
        return 1, rows, np.array([rows.shape[0]], dtype=np.int32), None
    if kind == "knots":
        _, ns, ds, np_, dp, knots = mode_args
        return (2, np.ascontiguousarray(knots, dtype=np.float64), np.array([ns, np_], dtype=np.int32),
                np.array([ds, dp], dtype=np.float64))
    raise ValueError(kind)


def run18(flat, controls, n, nsteps, tire_nr=False, ovr=None, init=None, t0=0.0, dt=1e-3, every=0, threads=0):
    return run(18, flat, controls, n, nsteps, tire_nr, ovr, init, t0, dt, every, threads)


def record_len(dof, tire_nr=False):
    return getattr(lib(), f"orc{dof}_record_len")(int(tire_nr))


def init24(flat, records, tire_nr=False, ovr=None):
    """initializeTireSus (dof24.cpp:34-89) applied to records (n, rec), returned as a new array."""
    recs = np.array(records, dtype=np.float64, order="C", copy=True)
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    ovr = None if ovr is None else np.ascontiguousarray(ovr, dtype=np.float64)
    lib().orc24_init(int(tire_nr), flat.ctypes.data, _ptr(ovr), recs.shape[0], recs.ctypes.data)
    return recs


def run(dof, flat, controls, n, nsteps, tire_nr=False, ovr=None, init=None, t0=0.0, dt=1e-3, every=0, threads=0,
        horizon=None):
    """Integrates n vehicles of one parameter set of the 11-, 18- or 24-DOF model.

    controls: ("steps", (nsteps, 3, n)) | ("table", rows (m, 4)) | ("knots", ns, ds, np, dp, (ns+2np, n)).
    horizon (optional, (n,) ints): vehicle i stops after horizon[i] steps; later samples repeat its final
    record.  Returns (final records (n, rec), samples (nsteps // every, n, rec) or None)."""
    L = lib()
    rl = record_len(dof, tire_nr)
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    mode, ctrl, im, dm = _ctrl(controls)
    ovr = None if ovr is None else np.ascontiguousarray(ovr, dtype=np.float64)
    init = None if init is None else np.ascontiguousarray(np.broadcast_to(init, (n, rl)), dtype=np.float64)
    fin = np.zeros((n, rl))
    out = np.zeros((nsteps // every, n, rl)) if every > 0 else None
    if horizon is not None:
        hz = np.ascontiguousarray(horizon, dtype=np.int32)
        assert hz.shape == (n,)
        assert L.orc_run_h(dof, int(tire_nr), _ptr(flat), _ptr(ovr), _ptr(init), n, t0, dt, nsteps, _ptr(hz), mode,
                           _ptr(ctrl), _ptr(im), _ptr(dm), every, _ptr(out), _ptr(fin), threads) == 0
        return fin, out
    getattr(L, f"orc{dof}_run")(int(tire_nr), _ptr(flat), _ptr(ovr), _ptr(init), n, t0, dt, nsteps, mode,
                                 _ptr(ctrl), _ptr(im), _ptr(dm), every, _ptr(out), _ptr(fin), threads)
    return fin, out


def loss_tangents18(flat, ovr, init, ctrl, xyref, dctrl, dovr, dinit, dt=1e-3, tire_nr=False, threads=0):
    """Tracking loss of n 18-DOF vehicles and its forward-mode directional derivatives (dual-number
    oracle): ctrl / dctrl (T, 3, n), ovr / dovr (7, n), init / dinit (n, rec), xyref (T, 2, n).  Returns
    (loss (n,), dloss (3, n): along dctrl, dovr, dinit)."""
    c = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    flat, ovr, init, ctrl, xyref, dctrl, dovr, dinit = map(c, (flat, ovr, init, ctrl, xyref, dctrl, dovr, dinit))
    T, _, n = ctrl.shape
    assert xyref.shape == (T, 2, n) and dctrl.shape == ctrl.shape and dovr.shape == ovr.shape == (7, n)
    assert init.shape == dinit.shape == (n, record_len(18, tire_nr))
    loss, dloss = np.zeros(n), np.zeros((3, n))
    assert lib().orc18_loss_tangents(int(tire_nr), _ptr(flat), _ptr(ovr), _ptr(init), n, T, dt, _ptr(ctrl),
                                     _ptr(xyref), _ptr(dctrl), _ptr(dovr), _ptr(dinit), _ptr(loss), _ptr(dloss),
                                     threads) == 0
    return loss, dloss


def dual_op(op, x, y=0.0):
    """(value, d/dx, d/dy) of one dual-scalar operation (see orc_dual_op)."""
    dx, dy = ctypes.c_double(), ctypes.c_double()
    v = lib().orc_dual_op(op, x, y, ctypes.byref(dx), ctypes.byref(dy))
    return v, dx.value, dy.value


OP_NAMES = ("add", "mul", "div", "sqrt", "sin", "cos", "tan", "atan2", "hypot", "tanh")


def count_ops18(flat, controls, nsteps, tire_nr=False, dt=1e-3):
    """Operation tallies over nsteps of one vehicle (the oracle run with a counting scalar)."""
    L = lib()
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    mode, ctrl, im, dm = _ctrl(controls)
    c = np.zeros(len(OP_NAMES), dtype=np.int64)
    L.orc18_count_ops(int(tire_nr), _ptr(flat), None, 0.0, dt, nsteps, mode, _ptr(ctrl), _ptr(im), _ptr(dm),
                      _ptr(c))
    return dict(zip(OP_NAMES, c.tolist()))


# distinct-argument tallies: a libm call repeating one of the same step with the same argument is
# counted in the plain tally only
DISTINCT_NAMES = ("sin_distinct", "cos_distinct", "tan_distinct", "atan2_distinct", "hypot_distinct", "tanh_distinct")


def count_ops(dof, flat, controls, nsteps, tire_nr=False, init=None, dt=1e-3):
    """Operation tallies over nsteps of one vehicle of the 11-, 18- or 24-DOF model (OP_NAMES, then
    DISTINCT_NAMES)."""
    L = lib()
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    init = None if init is None else np.ascontiguousarray(init, dtype=np.float64)
    mode, ctrl, im, dm = _ctrl(controls)
    c = np.zeros(len(OP_NAMES) + len(DISTINCT_NAMES), dtype=np.int64)
    assert L.orc_count_ops(dof, int(tire_nr), _ptr(flat), _ptr(init), 0.0, dt, nsteps, mode, _ptr(ctrl), _ptr(im),
                           _ptr(dm), _ptr(c)) == 0
    return dict(zip(OP_NAMES + DISTINCT_NAMES, c.tolist()))


def count_veh_rhs18(flat, rec, fx, fy):
    """(add, mul, div, sqrt) of one computeVehRHS evaluation."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (flat, rec, fx, fy)]
    c = np.zeros(4, dtype=np.int64)
    lib().orc_count_veh_rhs18(*[x.ctypes.data for x in a], c.ctypes.data)
    return tuple(c.tolist())


def count_tmxy(s, df0, sm, fm, ss, fs):
    """(add, mul, div, sqrt) of one tmxy_combined evaluation."""
    c = np.zeros(4, dtype=np.int64)
    lib().orc_count_tmxy(s, df0, sm, fm, ss, fs, c.ctypes.data)
    return tuple(c.tolist())


def tmxy(s, df0, sm, fm, ss, fs):
    f, fos = ctypes.c_double(), ctypes.c_double()
    lib().orc_tmxy(s, df0, sm, fm, ss, fs, ctypes.byref(f), ctypes.byref(fos))
    return f.value, fos.value


def interp_l(fz, w1, w2, pn):
    return lib().orc_interp_l(fz, w1, w2, pn)


def interp_q(fz, w1, w2, pn):
    return lib().orc_interp_q(fz, w1, w2, pn)


def sine_step(x, x1, y1, x2, y2):
    return lib().orc_sine_step(x, x1, y1, x2, y2)


def map_y(xy, x):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    return lib().orc_map_y(xy.ctypes.data, xy.shape[0], x)


def diff_split(torque, sl, sr):
    a, b = ctypes.c_double(), ctypes.c_double()
    lib().orc_diff_split(torque, sl, sr, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def loads18(flat, rec, tire_nr=False):
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    out = np.zeros(4)
    lib().orc18_loads(int(tire_nr), flat.ctypes.data, rec.ctypes.data, out.ctypes.data)
    return out


def veh_rhs18(flat, rec, fx, fy, tire_nr=False):
    flat = np.ascontiguousarray(flat, dtype=np.float64)
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    fx = np.ascontiguousarray(fx, dtype=np.float64)
    fy = np.ascontiguousarray(fy, dtype=np.float64)
    out = np.zeros(4)
    lib().orc18_veh_rhs(int(tire_nr), flat.ctypes.data, rec.ctypes.data, fx.ctypes.data, fy.ctypes.data,
                        out.ctypes.data)
    return out
