"""MI355X batched half-implicit solver for the 11/18/24-DOF vehicle models (HIP, gfx950).

The compute path is the in-tree liblfdm_amd.so (HIP kernels behind the C ABI of
include/lfdm_amd.h); this package is a thin host-side mirror of the reference solver API.
"""
from ._lib import LIB_PATH, LfdmError
from .solver import (ARITH_FAST, ARITH_STRICT, LAYOUT_BY_HORIZON, LAYOUT_GROUPED, OUTPUT_COLUMNS, OVERRIDE_FIELDS, Fleet, SolverHalfImplicitGPU,
                     TMeasy, TMeasyNr, load_driver_file, params_from_json, record_fields, record_index, record_len)

__all__ = [
    "ARITH_FAST",
    "ARITH_STRICT",
    "Fleet",
    "LAYOUT_BY_HORIZON",
    "LAYOUT_GROUPED",
    "LIB_PATH",
    "LfdmError",
    "OUTPUT_COLUMNS",
    "OVERRIDE_FIELDS",
    "SolverHalfImplicitGPU",
    "TMeasy",
    "TMeasyNr",
    "load_driver_file",
    "params_from_json",
    "record_fields",
    "record_index",
    "record_len",
]
