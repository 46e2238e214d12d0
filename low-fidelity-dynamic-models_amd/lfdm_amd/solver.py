"""Python mirror of the reference GPU solver classes over the C ABI.

``SolverHalfImplicitGPU(dof, total)`` follows d18SolverHalfImplicitGPU
(wheeled_vehicle_models/18dof-gpu/dof18_halfImplicit_gpu.cuh:15-343) and, for dof 11 and 24,
d11SolverHalfImplicitGPU / d24SolverHalfImplicitGPU: Construct, Initialize,
the Set*/Get* methods, SetOutput, Solve, SolveStep and GetSimState keep the reference's names,
argument order and meaning.  States are exchanged as full records (numpy float64 arrays in the
order documented in include/lfdm_amd.h); batched methods move whole fleets at once.
"""
import ctypes

import numpy as np

from ._lib import (LFDM_ARITH_FAST, LFDM_ARITH_STRICT, LFDM_LAYOUT_BY_HORIZON, LFDM_LAYOUT_GROUPED, LFDM_TIRE_TMEASY,
                   LFDM_TIRE_TMEASYNR, check, enc, lib)

TMeasy = LFDM_TIRE_TMEASY
TMeasyNr = LFDM_TIRE_TMEASYNR
LAYOUT_GROUPED = LFDM_LAYOUT_GROUPED
LAYOUT_BY_HORIZON = LFDM_LAYOUT_BY_HORIZON
ARITH_FAST = LFDM_ARITH_FAST
ARITH_STRICT = LFDM_ARITH_STRICT

# Field names of the records (include/lfdm_amd.h).
TIRE_FIELDS = ("xe", "ye", "xedot", "yedot", "omega", "dOmega", "xt", "rStat", "fx", "fy", "fz", "vsx", "vsy",
               "My", "engTor")
TIRE_FIELDS_NR = TIRE_FIELDS[4:]
VEH_FIELDS = ("x", "y", "u", "v", "psi", "wz", "phi", "wx", "udot", "vdot", "wxdot", "wzdot", "crankOmega",
              "dOmega_crank", "current_gr")
TIRES = ("lf", "rf", "lr", "rr")
VEH_FIELDS_11 = ("x", "y", "u", "v", "psi", "wz", "udot", "vdot", "wzdot", "crankOmega", "dOmega_crank", "current_gr")
TIRES_11 = ("f", "r")
TIRE_FIELDS_24 = ("xe", "ye", "xedot", "yedot", "omega", "dOmega", "xt", "dxt", "rStat", "fx", "fy", "fz", "fxg",
                  "fyg", "fzg", "vgx", "vgy", "vgz", "vsx", "vsy", "vsz", "My", "engTor")
TIRE_FIELDS_24_NR = TIRE_FIELDS_24[4:]
SUS_FIELDS = ("xs", "xsi", "dxs", "ls", "lsi", "us", "vs", "ws", "dus", "dvs", "dws", "uu", "vu", "wu", "duu", "dvu",
              "dwu", "fxs", "fys", "fzs", "fzd", "mx", "my", "mz")
VEH_FIELDS_24 = ("x", "y", "z", "u", "v", "w", "psi", "wz", "phi", "wx", "theta", "wy", "udot", "vdot", "wdot",
                 "wxdot", "wydot", "wzdot", "dtheta", "dphi", "dpsi", "crankOmega", "dOmega_crank", "current_gr")
OVERRIDE_FIELDS = ("m", "jz", "jx", "dfy0Pn", "dfy0P2n", "torque_scale", "maxBrakeTorque")
# Output channels of SetOutput / Solve (the reference's device response) and their CSV columns.
OUTPUT_COLUMNS = {
    11: ("time", "x", "y", "vx", "vy", "yaw", "yaw_rate", "wf", "wr"),
    18: ("time", "x", "y", "vx", "vy", "roll", "yaw", "roll_rate", "yaw_rate", "wlf", "wrf", "wlr", "wrr"),
    24: ("time", "x", "y", "vx", "vy", "roll", "yaw", "pitch", "roll_rate", "yaw_rate", "pitch_rate", "wlf", "wrf",
         "wlr", "wrr"),
}


def record_fields(dof=18, tire_type=TMeasy):
    """Names of the record entries, e.g. 'lf.omega', 'rr.fzd' (24-DOF strut) or 'x'."""
    nr = tire_type == TMeasyNr
    if dof == 18:
        tf = TIRE_FIELDS_NR if nr else TIRE_FIELDS
        return [f"{t}.{f}" for t in TIRES for f in tf] + list(VEH_FIELDS)
    if dof == 11:
        tf = TIRE_FIELDS_NR if nr else TIRE_FIELDS
        return [f"{t}.{f}" for t in TIRES_11 for f in tf] + list(VEH_FIELDS_11)
    if dof == 24:
        tf = TIRE_FIELDS_24_NR if nr else TIRE_FIELDS_24
        return ([f"{t}.{f}" for t in TIRES for f in tf] + [f"sus_{t}.{f}" for t in TIRES for f in SUS_FIELDS] +
                list(VEH_FIELDS_24))
    raise ValueError(f"unsupported model: {dof}-DOF")


def record_index(name, dof=18, tire_type=TMeasy):
    return record_fields(dof, tire_type).index(name)


def record_len(dof=18, tire_type=TMeasy):
    return check(lib.lfdm_record_len(dof, tire_type))


def params_from_json(dof, veh_json, tire_json, sus_json=None, tire_type=TMeasy):
    """Flat parameter vector (setVehParamsJSON + setTireParamsJSON + tireInit semantics)."""
    buf = np.zeros(8192)
    n = ctypes.c_int(0)
    check(lib.lfdm_params_from_json(dof, enc(veh_json), enc(tire_json), enc(sus_json), tire_type,
                                    buf.ctypes.data, buf.size, ctypes.byref(n)))
    return buf[: n.value].copy()


def load_driver_file(path):
    """Driver table rows (time, steering, throttle, braking) as an (n, 4) array."""
    cap = 1 << 16
    rows = np.zeros((cap, 4))
    n = ctypes.c_int(0)
    check(lib.lfdm_driver_file_load(enc(path), rows.ctypes.data, cap, ctypes.byref(n)))
    return rows[: n.value].copy()


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class SolverHalfImplicitGPU:
    """One solver handle; see the module docstring."""

    def __init__(self, dof, total_num_vehicles):
        h = ctypes.c_void_p()
        check(lib.lfdm_create(dof, total_num_vehicles, ctypes.byref(h)))
        self._h = h
        self.dof = dof
        self.total = total_num_vehicles
        self.tire_type = None

    def close(self):
        if self._h:
            lib.lfdm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- construction ------------------------------------------------------------------
    def Construct(self, veh_params_file, tire_params_file, *args, **kw):
        """Construct(veh, tire, n[, driver_file][, tire_type]) for 11/18-DOF and
        Construct(veh, tire, sus, n[, driver_file][, tire_type]) for 24-DOF, as in the reference."""
        names = (["sus_params_file"] if self.dof == 24 else []) + ["num_vehicles", "driver_file", "tire_type"]
        a = dict(zip(names, args))
        a.update(kw)
        tire_type = a.get("tire_type", TMeasy)
        check(lib.lfdm_construct(self._h, enc(veh_params_file), enc(tire_params_file), enc(a.get("sus_params_file")),
                                 a["num_vehicles"], enc(a.get("driver_file")), tire_type))
        self.tire_type = tire_type

    def construct_flat(self, flat, num_vehicles, driver_rows=None, tire_type=TMeasy):
        flat = _f64(flat)
        rows = None if driver_rows is None else _f64(driver_rows)
        check(lib.lfdm_construct_flat(self._h, flat.ctypes.data, flat.size, tire_type, num_vehicles,
                                      None if rows is None else rows.ctypes.data,
                                      0 if rows is None else rows.shape[0]))
        self.tire_type = tire_type

    def Initialize(self, record, num_vehicles):
        rec = _f64(record)
        check(lib.lfdm_initialize(self._h, rec.ctypes.data, num_vehicles))

    def initialize_batch(self, records):
        """Appends len(records) vehicles, each with its own record (rows of an (n, rec) array)."""
        rec = _f64(records)
        check(lib.lfdm_initialize_batch(self._h, rec.ctypes.data, rec.shape[0]))

    def zero_record(self):
        return np.zeros(record_len(self.dof, self.tire_type or TMeasy))

    # -- settings ----------------------------------------------------------------------
    def SetEndTime(self, tend):
        check(lib.lfdm_set_end_time(self._h, tend))

    def SetTimeStep(self, step):
        check(lib.lfdm_set_time_step(self._h, step))

    def SetKernelSimTime(self, t):
        check(lib.lfdm_set_kernel_sim_time(self._h, t))

    def SetHostDumpTime(self, t):
        check(lib.lfdm_set_host_dump_time(self._h, t))

    def SetThreadsPerBlock(self, n):
        check(lib.lfdm_set_threads_per_block(self._h, n))

    def GetStep(self):
        v = ctypes.c_double()
        check(lib.lfdm_get_step(self._h, ctypes.byref(v)))
        return v.value

    def GetEndTime(self):
        v = ctypes.c_double()
        check(lib.lfdm_get_end_time(self._h, ctypes.byref(v)))
        return v.value

    def GetTireType(self):
        v = ctypes.c_int()
        check(lib.lfdm_get_tire_type(self._h, ctypes.byref(v)))
        return v.value

    def SetOutput(self, output_file, output_freq, store_all=False, no_outs=50):
        check(lib.lfdm_set_output(self._h, enc(output_file), output_freq, int(store_all), no_outs))

    # -- integration -------------------------------------------------------------------
    def Solve(self):
        check(lib.lfdm_solve(self._h))

    def SolveStep(self, t, steering, throttle, braking):
        out = ctypes.c_double()
        check(lib.lfdm_solve_step(self._h, t, steering, throttle, braking, ctypes.byref(out)))
        return out.value

    def GetSimState(self, vehicle_index):
        rec = self.zero_record()
        check(lib.lfdm_get_sim_state(self._h, vehicle_index, rec.ctypes.data))
        return rec

    # -- batched extensions -------------------------------------------------------------
    def get_states(self):
        out = np.zeros((self.total, record_len(self.dof, self.tire_type or TMeasy)))
        check(lib.lfdm_get_states(self._h, out.ctypes.data))
        return out

    def set_states(self, records):
        rec = _f64(records)
        check(lib.lfdm_set_states(self._h, rec.ctypes.data))

    def set_overrides(self, ovr):
        """ovr: (7, total) array of m, jz, jx, dfy0Pn, dfy0P2n, torque_scale, maxBrakeTorque."""
        ovr = _f64(ovr)
        assert ovr.shape == (len(OVERRIDE_FIELDS), self.total)
        check(lib.lfdm_set_overrides(self._h, ovr.ctypes.data))

    def get_vehicle_params(self, vehicle):
        """The vehicle's full flat parameter vector (its set with its per-vehicle parameters applied)."""
        buf = np.zeros(8192)
        n = ctypes.c_int(0)
        check(lib.lfdm_get_vehicle_params(self._h, vehicle, buf.ctypes.data, buf.size, ctypes.byref(n)))
        return buf[: n.value].copy()

    def set_vehicle_params(self, vehicle, flat):
        """Gives one vehicle this flat parameter vector (any field), as editing SimData of the reference."""
        f = _f64(flat)
        check(lib.lfdm_set_vehicle_params(self._h, vehicle, f.ctypes.data, f.size))

    def set_overrides_device(self, device_ptr):
        """Per-vehicle parameters from device memory, (7, total) float64 (e.g. a torch tensor's data_ptr())."""
        check(lib.lfdm_set_overrides_device(self._h, ctypes.c_void_p(device_ptr)))

    def set_states_device(self, device_ptr):
        """Records from device memory, (total, record_len) float64."""
        check(lib.lfdm_set_states_device(self._h, ctypes.c_void_p(device_ptr)))

    def set_controls_knots(self, ns, ds, np_, dp, knots):
        knots = _f64(knots)
        assert knots.shape == (ns + 2 * np_, self.total)
        check(lib.lfdm_set_controls_knots(self._h, ns, ds, np_, dp, knots.ctypes.data))

    def set_controls_steps(self, ctrl):
        """ctrl: (nsteps, 3, total) steering, throttle, braking per step (host array)."""
        ctrl = _f64(ctrl)
        assert ctrl.ndim == 3 and ctrl.shape[1:] == (3, self.total)
        self._steps_keep = ctrl
        check(lib.lfdm_set_controls_steps(self._h, ctrl.shape[0], ctrl.ctypes.data))

    def set_controls_steps_device(self, nsteps, device_ptr):
        check(lib.lfdm_set_controls_steps_device(self._h, nsteps, ctypes.c_void_p(device_ptr)))

    def set_controls_table(self):
        check(lib.lfdm_set_controls_table(self._h))

    def advance(self, nsteps):
        check(lib.lfdm_advance(self._h, nsteps))

    def set_time(self, t):
        check(lib.lfdm_set_time(self._h, t))

    # -- batching form -----------------------------------------------------------------------
    def set_horizons(self, steps):
        """Per-vehicle step counts from t = 0 after which each vehicle stops (None: no horizons)."""
        if steps is None:
            check(lib.lfdm_set_horizons(self._h, None))
            return
        a = np.ascontiguousarray(steps, dtype=np.int32)
        if a.shape != (self.total,):
            raise ValueError(f"horizons must have shape ({self.total},)")
        check(lib.lfdm_set_horizons(self._h, a.ctypes.data))

    def set_layout(self, layout):
        check(lib.lfdm_set_layout(self._h, layout))

    def get_layout(self):
        v = ctypes.c_int()
        check(lib.lfdm_get_layout(self._h, ctypes.byref(v)))
        return v.value

    def set_arithmetic(self, arith):
        """ARITH_FAST (default) or ARITH_STRICT: the step with the reference's own quotients,
        atan2 / tan and hypot (see lfdm_set_arithmetic)."""
        check(lib.lfdm_set_arithmetic(self._h, arith))

    def get_arithmetic(self):
        v = ctypes.c_int()
        check(lib.lfdm_get_arithmetic(self._h, ctypes.byref(v)))
        return v.value

    # -- differentiation form (device pointers, e.g. torch tensors' data_ptr()) ---------------
    def adjoint_forward(self, nsteps, checkpoint_every, ctrl_ptr, xy_ptr):
        check(lib.lfdm_adjoint_forward(self._h, nsteps, checkpoint_every, ctypes.c_void_p(ctrl_ptr),
                                       ctypes.c_void_p(xy_ptr)))

    def adjoint_backward(self, xy_bar_ptr, ctrl_bar_ptr, state_bar_ptr, param_bar_ptr):
        check(lib.lfdm_adjoint_backward(self._h, ctypes.c_void_p(xy_bar_ptr), ctypes.c_void_p(ctrl_bar_ptr),
                                        ctypes.c_void_p(state_bar_ptr), ctypes.c_void_p(param_bar_ptr)))

    def adjoint_set_tape_budget(self, nbytes):
        """Device bytes the adjoint's tape may take from the next forward on; 0 (the default): half of
        the device memory free at the forward."""
        check(lib.lfdm_adjoint_set_tape_budget(self._h, int(nbytes)))

    def adjoint_tape_bytes(self):
        """(bytes of tape, segments per backward group) of the last adjoint forward."""
        b, g = ctypes.c_ulonglong(), ctypes.c_int()
        check(lib.lfdm_adjoint_tape_bytes(self._h, ctypes.byref(b), ctypes.byref(g)))
        return b.value, g.value

    def adjoint_generation(self):
        v = ctypes.c_ulonglong()
        check(lib.lfdm_adjoint_generation(self._h, ctypes.byref(v)))
        return v.value

    def step_jacobian(self, ctrl_ptr, jac_ptr):
        """Exact Jacobian of one step at every vehicle's current record (device pointers: controls
        [3][total], Jacobian [21][24][total]; see lfdm_step_jacobian).  Returns (rows, cols)."""
        r, c = ctypes.c_int(), ctypes.c_int()
        check(lib.lfdm_step_jacobian(self._h, ctypes.c_void_p(ctrl_ptr), ctypes.c_void_p(jac_ptr), ctypes.byref(r),
                                     ctypes.byref(c)))
        return r.value, c.value

    def set_stream(self, stream_ptr):
        check(lib.lfdm_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def synchronize(self):
        check(lib.lfdm_synchronize(self._h))

    def last_launch_ms(self):
        v = ctypes.c_double()
        check(lib.lfdm_last_launch_ms(self._h, ctypes.byref(v)))
        return v.value


class Fleet:
    """Mixed fleet over solver handles (one bucket per handle, normally one per model), run as one
    HIP graph per lfdm_fleet_run.  The solvers must outlive the fleet."""

    def __init__(self, solvers=()):
        f = ctypes.c_void_p()
        check(lib.lfdm_fleet_create(ctypes.byref(f)))
        self._f = f
        self.solvers = []
        for s in solvers:
            self.add(s)

    def add(self, solver):
        idx = check(lib.lfdm_fleet_add(self._f, solver._h))
        self.solvers.append(solver)
        return idx

    def close(self):
        if self._f:
            lib.lfdm_fleet_destroy(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_stream(self, stream_ptr):
        check(lib.lfdm_fleet_set_stream(self._f, ctypes.c_void_p(stream_ptr)))

    def set_config(self, bucket, block, steps_per_launch=0, waves_per_simd=0):
        check(lib.lfdm_fleet_set_config(self._f, bucket, block, steps_per_launch, waves_per_simd))

    def get_config(self, bucket):
        """(block, steps_per_launch, waves_per_simd, layout) of a bucket."""
        b, s, w, l = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib.lfdm_fleet_get_config(self._f, bucket, ctypes.byref(b), ctypes.byref(s), ctypes.byref(w),
                                        ctypes.byref(l)))
        return b.value, s.value, w.value, l.value

    def autotune(self, nsteps, trial_steps=1000):
        check(lib.lfdm_fleet_autotune(self._f, nsteps, trial_steps))
        return [self.get_config(b) for b in range(len(self.solvers))]

    def run(self, nsteps):
        check(lib.lfdm_fleet_run(self._f, nsteps))

    def synchronize(self):
        check(lib.lfdm_fleet_synchronize(self._f))

    def last_ms(self):
        v = ctypes.c_double()
        check(lib.lfdm_fleet_last_ms(self._f, ctypes.byref(v)))
        return v.value
