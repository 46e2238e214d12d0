"""ctypes binding of oracle/_ref/libref18.so: the reference CPU half-implicit solver compiled from
its own sources (wheeled_vehicle_models/18dof/*.cpp) plus the shim oracle/ref_shim18.cpp.

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


def lib18():
    if 18 not in _libs:
        L = ctypes.CDLL(os.path.join(REF_DIR, "libref18.so"))
        p, d, i, s = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
        L.ref18_create.restype = p
        L.ref18_create.argtypes = [s, s, i, s]
        L.ref18_destroy.argtypes = [p]
        L.ref18_set_step.argtypes = [p, d]
        L.ref18_set_param.argtypes = [p, s, d]
        L.ref18_set_param.restype = i
        L.ref18_dump_params.argtypes = [p, p]
        L.ref18_dump_params.restype = i
        L.ref18_get_record.argtypes = [p, p]
        L.ref18_set_record.argtypes = [p, p]
        L.ref18_run.argtypes = [p, d, i, p, i, p]
        L.ref18_run.restype = d
        L.ref18_solve.argtypes = [p]
        L.ref18_record_len.argtypes = [i]
        L.ref18_record_len.restype = i
        _libs[18] = L
    return _libs[18]


class Ref18:
    """One reference d18SolverHalfImplicit (constructed from JSON files, zero initial state)."""

    def __init__(self, veh, tire, tire_nr=False, driver=None, dt=1e-3):
        L = lib18()
        self.L = L
        self.nr = bool(tire_nr)
        self.h = L.ref18_create(os.fsencode(veh), os.fsencode(tire), int(tire_nr),
                                os.fsencode(driver) if driver else None)
        L.ref18_set_step(self.h, dt)
        self.rl = L.ref18_record_len(int(tire_nr))

    def close(self):
        if self.h:
            self.L.ref18_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def params(self):
        buf = np.zeros(8192)
        n = self.L.ref18_dump_params(self.h, buf.ctypes.data)
        return buf[:n].copy()

    def set_param(self, name, value):
        assert self.L.ref18_set_param(self.h, name.encode(), float(value)) == 0, name

    def record(self):
        r = np.zeros(self.rl)
        self.L.ref18_get_record(self.h, r.ctypes.data)
        return r

    def set_record(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self.L.ref18_set_record(self.h, r.ctypes.data)

    def run(self, nsteps, ctrl=None, every=0, t0=0.0):
        """ctrl: (nsteps, 3) steering, throttle, braking or None for the driver file."""
        c = None if ctrl is None else np.ascontiguousarray(ctrl, dtype=np.float64)
        out = np.zeros((nsteps // every, self.rl)) if every > 0 else None
        t = self.L.ref18_run(self.h, t0, nsteps, None if c is None else c.ctypes.data, every,
                             None if out is None else out.ctypes.data)
        return t, out

    def solve(self):
        self.L.ref18_solve(self.h)
