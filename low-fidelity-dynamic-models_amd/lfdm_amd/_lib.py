"""ctypes binding of the C ABI declared in include/lfdm_amd.h.

The shared library is built in-tree (``make -C low-fidelity-dynamic-models_amd``); importing this
module fails loudly if it is missing, since there is no other implementation of the solver.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LFDM_AMD_LIB") or os.path.join(_HERE, "liblfdm_amd.so")

LFDM_OK = 0
LFDM_TIRE_TMEASY = 0
LFDM_TIRE_TMEASYNR = 1
LFDM_LAYOUT_GROUPED = 0
LFDM_LAYOUT_BY_HORIZON = 1
LFDM_ARITH_FAST = 0
LFDM_ARITH_STRICT = 1
ERRORS = {
    -1: "LFDM_E_ARG",
    -2: "LFDM_E_IO",
    -3: "LFDM_E_PARSE",
    -4: "LFDM_E_STATE",
    -5: "LFDM_E_CAPACITY",
    -6: "LFDM_E_HIP",
    -7: "LFDM_E_UNSUPPORTED",
}


class LfdmError(RuntimeError):
    """A C-ABI call returned an error status."""

    def __init__(self, code, message):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code
        self.name = ERRORS.get(code, str(code))


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is not built; run `make -C low-fidelity-dynamic-models_amd` "
        "(or __graft_entry__.build()) first"
    )

# One HIP runtime per process: torch ships its own libamdhip64 / libhsa-runtime64, and a process that
# loads this library first and torch afterwards ends up with two runtimes whose device memory and
# streams are not shared (torch tensors handed to the solver's kernels are then invalid).  Importing
# torch first, when it is installed, makes the library bind to the runtime torch already loaded.
try:
    import torch  # noqa: F401
except ImportError:
    pass

lib = ctypes.CDLL(LIB_PATH)

_p = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
_u = ctypes.c_uint
_s = ctypes.c_char_p
_pd = ctypes.POINTER(ctypes.c_double)

# name -> argument types (every function returns int)
SIGNATURES = {
    "lfdm_record_len": [_i, _i],
    "lfdm_params_from_json": [_i, _s, _s, _s, _i, _p, _i, ctypes.POINTER(_i)],
    "lfdm_driver_file_load": [_s, _p, _i, ctypes.POINTER(_i)],
    "lfdm_create": [_i, _u, ctypes.POINTER(_p)],
    "lfdm_destroy": [_p],
    "lfdm_construct": [_p, _s, _s, _s, _u, _s, _i],
    "lfdm_construct_flat": [_p, _p, _i, _i, _u, _p, _i],
    "lfdm_initialize": [_p, _p, _u],
    "lfdm_initialize_batch": [_p, _p, _u],
    "lfdm_set_end_time": [_p, _d],
    "lfdm_set_time_step": [_p, _d],
    "lfdm_set_kernel_sim_time": [_p, _d],
    "lfdm_set_host_dump_time": [_p, _d],
    "lfdm_set_threads_per_block": [_p, _u],
    "lfdm_get_step": [_p, _pd],
    "lfdm_get_end_time": [_p, _pd],
    "lfdm_get_kernel_sim_time": [_p, _pd],
    "lfdm_get_tire_type": [_p, ctypes.POINTER(_i)],
    "lfdm_set_output": [_p, _s, _d, _i, _u],
    "lfdm_solve": [_p],
    "lfdm_solve_step": [_p, _d, _d, _d, _d, _pd],
    "lfdm_get_sim_state": [_p, _u, _p],
    "lfdm_get_states": [_p, _p],
    "lfdm_set_states": [_p, _p],
    "lfdm_set_overrides": [_p, _p],
    "lfdm_set_overrides_device": [_p, _p],
    "lfdm_get_vehicle_params": [_p, _u, _p, _i, ctypes.POINTER(_i)],
    "lfdm_set_vehicle_params": [_p, _u, _p, _i],
    "lfdm_set_states_device": [_p, _p],
    "lfdm_adjoint_generation": [_p, ctypes.POINTER(ctypes.c_ulonglong)],
    "lfdm_set_controls_knots": [_p, _i, _d, _i, _d, _p],
    "lfdm_set_controls_steps": [_p, _i, _p],
    "lfdm_set_controls_steps_device": [_p, _i, _p],
    "lfdm_set_controls_table": [_p],
    "lfdm_set_controls_shared_device": [_p, _i, _p],
    "lfdm_chunk_steps": [_p, _d, ctypes.POINTER(_i)],
    "lfdm_initialized_count": [_p, ctypes.POINTER(_u)],
    "lfdm_advance": [_p, _i],
    "lfdm_set_time": [_p, _d],
    "lfdm_adjoint_forward": [_p, _i, _i, _p, _p],
    "lfdm_adjoint_backward": [_p, _p, _p, _p, _p],
    "lfdm_adjoint_set_tape_budget": [_p, ctypes.c_ulonglong],
    "lfdm_adjoint_tape_bytes": [_p, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_int)],
    "lfdm_step_jacobian": [_p, _p, _p, _p, _p],
    "lfdm_set_stream": [_p, _p],
    "lfdm_synchronize": [_p],
    "lfdm_last_launch_ms": [_p, _pd],
    "lfdm_set_horizons": [_p, _p],
    "lfdm_set_layout": [_p, _i],
    "lfdm_get_layout": [_p, ctypes.POINTER(_i)],
    "lfdm_set_arithmetic": [_p, _i],
    "lfdm_get_arithmetic": [_p, ctypes.POINTER(_i)],
    "lfdm_fleet_create": [ctypes.POINTER(_p)],
    "lfdm_fleet_destroy": [_p],
    "lfdm_fleet_add": [_p, _p],
    "lfdm_fleet_set_stream": [_p, _p],
    "lfdm_fleet_set_config": [_p, _i, _i, _i, _i],
    "lfdm_fleet_get_config": [_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                              ctypes.POINTER(_i)],
    "lfdm_fleet_autotune": [_p, _i, _i],
    "lfdm_fleet_run": [_p, _i],
    "lfdm_fleet_synchronize": [_p],
    "lfdm_fleet_last_ms": [_p, _pd],
    "lfdm_integrate_noalloc": [_i, _p, _i, _i, _u, _p, _p, _p, _i, _d, _i, _p, _p],
}

for _name, _args in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = ctypes.c_int

lib.lfdm_last_error.argtypes = []
lib.lfdm_last_error.restype = ctypes.c_char_p


def check(rc):
    """Raises LfdmError for a negative status."""
    if rc < 0:
        raise LfdmError(rc, lib.lfdm_last_error().decode())
    return rc


def enc(s):
    return None if s is None else os.fsencode(s)
