"""ctypes binding of oracle/_ref/libref{11,18,24}.so: the reference CPU half-implicit solvers compiled
from their own sources (wheeled_vehicle_models/{11,18,24}dof/*.cpp) plus the shims oracle/ref_shim*.cpp.

Present only where oracle/_ref was built (make -C oracle ref, which needs the reference sources);
tests that use it skip when it is absent.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


def available(dof=18):
    return os.path.exists(os.path.join(REF_DIR, f"libref{dof}.so"))


_libs = {}


def _lib(dof):
    if dof not in _libs:
        L = ctypes.CDLL(os.path.join(REF_DIR, f"libref{dof}.so"))
        p, d, i, s = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        f = lambda name: getattr(L, f"ref{dof}_{name}")  # noqa: E731
        f("create").restype = p
        f("create").argtypes = [s, s, s, i, s] if dof == 24 else [s, s, i, s]
        f("destroy").argtypes = [p]
        f("set_step").argtypes = [p, d]
        f("set_param").argtypes = [p, s, d]
        f("set_param").restype = i
        f("dump_params").argtypes = [p, p]
        f("dump_params").restype = i
        f("get_record").argtypes = [p, p]
        f("run").argtypes = [p, d, i, p, i, p]
        f("run").restype = d
        f("solve").argtypes = [p]
        f("record_len").argtypes = [i]
        f("record_len").restype = i
        if dof == 18:
            f("set_record").argtypes = [p, p]
            f("step_jacobian").argtypes = [p, d, d, d, d, p, p]
            f("step_jacobian").restype = i
        if dof == 24:
            f("initialize").argtypes = [p, d]
        _libs[dof] = L
    return _libs[dof]


def lib18():
    return _lib(18)


class RefSolver:
    """One reference d{11,18,24}SolverHalfImplicit constructed from JSON files, zero initial state
    (24-DOF: Initialize with zero states and forward speed u0, which runs initializeTireSus).
    overrides: {name: value} set on the parsed parameters before Initialize."""

    def __init__(self, dof, veh, tire, tire_nr=False, driver=None, dt=1e-3, sus=None, u0=0.0, overrides=None):
        L = _lib(dof)
        self.L, self.dof = L, dof
        self._f = lambda name: getattr(L, f"ref{dof}_{name}")  # noqa: E731
        self.nr = bool(tire_nr)
        drv = os.fsencode(driver) if driver else None
        if dof == 24:
            self.h = self._f("create")(os.fsencode(veh), os.fsencode(tire), os.fsencode(sus), int(tire_nr), drv)
        else:
            self.h = self._f("create")(os.fsencode(veh), os.fsencode(tire), int(tire_nr), drv)
        self._f("set_step")(self.h, dt)
        for k, v in (overrides or {}).items():
            self.set_param(k, v)
        if dof == 24:
            self._f("initialize")(self.h, float(u0))
        self.rl = self._f("record_len")(int(tire_nr))

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        buf = np.zeros(8192)
        n = self._f("dump_params")(self.h, buf.ctypes.data)
        return buf[:n].copy()

    def set_param(self, name, value):
        assert self._f("set_param")(self.h, name.encode(), float(value)) == 0, name

    def record(self):
        r = np.zeros(self.rl)
        self._f("get_record")(self.h, r.ctypes.data)
        return r

    def set_record(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._f("set_record")(self.h, r.ctypes.data)

    def run(self, nsteps, ctrl=None, every=0, t0=0.0):
        """ctrl: (nsteps, 3) steering, throttle, braking or None for the driver file."""
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float64)
        out = np.zeros((nsteps // every, self.rl)) if every > 0 else None
        t = self._f("run")(self.h, t0, nsteps, None if c is None else c.ctypes.data, every,
                           None if out is None else out.ctypes.data)
        return t, out

    def solve(self):
        self._f("solve")(self.h)


class Ref18(RefSolver):
    def __init__(self, veh, tire, tire_nr=False, driver=None, dt=1e-3):
        super().__init__(18, veh, tire, tire_nr, driver, dt)

    def step_jacobian(self, t, steering, throttle, braking):
        """IntegrateStepWithJacobian from the current state: (J_state [n, n], J_controls [n, 2]) of the
        right-hand side in the reference's packY order; the solver then stands one step later."""
        js, jc = np.zeros(21 * 21), np.zeros(21 * 2)
        n = self._f("step_jacobian")(self.h, float(t), float(throttle), float(steering), float(braking),
                                     js.ctypes.data, jc.ctypes.data)
        return js[: n * n].reshape(n, n).copy(), jc[: 2 * n].reshape(n, 2).copy()
