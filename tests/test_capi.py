"""Host-side checks of the C ABI library (no GPU needed): exports, parameter parsing parity with the
reference's parser, driver-file loading, argument and ordering errors."""
import ctypes
import os
import re

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import LfdmError, SolverHalfImplicitGPU, TMeasy, TMeasyNr
from lfdm_amd._lib import lib
from cases18 import CASES, paths
from cases1124 import CASES as CASES1124, paths as paths1124
from conftest import DATA, ROOT

HEADER = os.path.join(ROOT, "include", "lfdm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int)\s+(lfdm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and nothing else is exported under the lfdm_ prefix that the header does not declare
    out = os.popen(f"nm -D --defined-only {lfdm_amd.LIB_PATH}").read()
    exported = set(re.findall(r"\b(lfdm_\w+)\b", out))
    assert exported == set(names), exported ^ set(names)


@pytest.mark.parametrize("name", sorted(CASES))
def test_params_match_reference_parser(name, golden18):
    veh, tire, nr, _, _ = paths(name)
    flat = lfdm_amd.params_from_json(18, veh, tire, tire_type=TMeasyNr if nr else TMeasy)
    np.testing.assert_array_equal(flat, golden18[f"{name}/params"])


@pytest.mark.parametrize("dof,name", [(d, n) for d in (11, 24) for n in sorted(CASES1124[d])])
def test_params_match_reference_parser_11_24(dof, name, request):
    """The 11-DOF loaders double the axle tire's stiffness, damping, inertia and force parameters
    (dof11.cpp:1036-1380); the 24-DOF ones append the suspension file (dof24.cpp:2192-2216)."""
    g = request.getfixturevalue(f"golden{dof}")
    veh, tire, nr, _, _, sus = paths1124(dof, name)
    flat = lfdm_amd.params_from_json(dof, veh, tire, sus, tire_type=TMeasyNr if nr else TMeasy)
    np.testing.assert_array_equal(flat, g[f"{name}/params"])


def test_record_layout():
    assert lfdm_amd.record_len(18, TMeasy) == 75
    assert lfdm_amd.record_len(18, TMeasyNr) == 59
    f = lfdm_amd.record_fields(18, TMeasy)
    assert f[4] == "lf.omega" and f[60] == "x" and f[-1] == "current_gr"
    for dof, nr, n in ((11, TMeasy, 42), (11, TMeasyNr, 34), (24, TMeasy, 212), (24, TMeasyNr, 196)):
        assert lfdm_amd.record_len(dof, nr) == n == len(lfdm_amd.record_fields(dof, nr))
    f24 = lfdm_amd.record_fields(24, TMeasy)
    assert f24[4] == "lf.omega" and f24[92] == "sus_lf.xs" and f24[-24] == "x" and f24[-1] == "current_gr"
    with pytest.raises(LfdmError):
        lfdm_amd.record_len(7, TMeasy)


def test_driver_file_loader_matches_text():
    path = os.path.join(DATA, "input", "acc3.txt")
    rows = lfdm_amd.load_driver_file(path)
    np.testing.assert_array_equal(rows, np.loadtxt(path, ndmin=2)[:, :4])


def test_parse_errors_are_reported():
    d = os.path.join(DATA, "18dof")
    # the bundled HMMWV TMeasyNr file spells its class key "VehicleType"; the reference aborts on it
    with pytest.raises(LfdmError) as e:
        lfdm_amd.params_from_json(18, f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasyNr.json", tire_type=TMeasyNr)
    assert e.value.name == "LFDM_E_PARSE" and "vehicleType" in str(e.value)
    with pytest.raises(LfdmError) as e:
        lfdm_amd.params_from_json(18, f"{d}/nope.json", f"{d}/HMMWV/tmeasy.json")
    assert e.value.name == "LFDM_E_IO"
    with pytest.raises(LfdmError) as e:
        lfdm_amd.load_driver_file(f"{d}/HMMWV/vehicle.json")  # JSON is no driver table
    assert e.value.name == "LFDM_E_IO"
    # the bundled 24-DOF Sedan vehicle_RWD.json is an 18-DOF file (no "jy"): the reference aborts
    d24 = os.path.join(DATA, "24dof", "Sedan")
    with pytest.raises(LfdmError) as e:
        lfdm_amd.params_from_json(24, f"{d24}/vehicle_RWD.json", f"{d24}/tmeasy.json", f"{d24}/suspension.json")
    assert e.value.name == "LFDM_E_PARSE" and "jy" in str(e.value)
    with pytest.raises(LfdmError) as e:  # the 24-DOF model needs its suspension file
        lfdm_amd.params_from_json(24, f"{d24}/vehicle.json", f"{d24}/tmeasy.json")
    assert e.value.name == "LFDM_E_ARG"


def test_handle_lifecycle_errors(tmp_path):
    d = os.path.join(DATA, "18dof")
    veh, tire = f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json"
    drv = os.path.join(DATA, "input", "acc3.txt")
    s = SolverHalfImplicitGPU(18, 10)
    s.Construct(veh, tire, 6, drv)
    with pytest.raises(LfdmError) as e:  # more vehicles than the handle holds
        s.Construct(veh, tire, 5, drv)
    assert e.value.name == "LFDM_E_CAPACITY"
    with pytest.raises(LfdmError) as e:  # tire types cannot be mixed within one solver
        s.Construct(f"{d}/Sedan/vehicle.json", f"{d}/Sedan/tmeasyNr.json", 2, drv, TMeasyNr)
    assert e.value.name == "LFDM_E_ARG"
    s.Construct(veh, tire, 4, drv)
    rec = s.zero_record()
    s.Initialize(rec, 10)
    with pytest.raises(LfdmError) as e:  # Solve before SetEndTime (the reference asserts)
        s.Solve()
    assert e.value.name == "LFDM_E_STATE"
    bad = rec.copy()
    bad[60] = np.nan
    s2 = SolverHalfImplicitGPU(18, 2)
    s2.Construct(veh, tire, 2, drv)
    with pytest.raises(LfdmError):
        s2.Initialize(bad, 2)
    # states are readable before the first launch
    np.testing.assert_array_equal(s.GetSimState(3), rec)
    with pytest.raises(LfdmError):
        s.GetSimState(10)
    s.SetOutput(str(tmp_path / "out"), 100)
    with pytest.raises(LfdmError) as e:  # SolveStep with CSV output (the reference exits)
        s.SolveStep(0.0, 0.0, 0.5, 0.0)
    assert e.value.name == "LFDM_E_STATE"
    assert s.GetStep() == 1e-3 and s.GetTireType() == TMeasy
    s.SetTimeStep(5e-4)
    s.SetEndTime(3.0)
    assert s.GetStep() == 5e-4 and s.GetEndTime() == 3.0
    for bad_call in (lambda: s.SetTimeStep(0.0), lambda: s.SetKernelSimTime(-1), lambda: s.SetThreadsPerBlock(0)):
        with pytest.raises(LfdmError):
            bad_call()
    s.adjoint_set_tape_budget(0)  # automatic (the default)
    with pytest.raises(LfdmError) as e:  # no adjoint forward yet
        s.adjoint_tape_bytes()
    assert e.value.name == "LFDM_E_STATE"
    s.close()
    s2.close()


def test_errors_without_gpu_are_loud():
    # with no GPU the first device use must fail with LFDM_E_HIP, never fall back to the CPU
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("GPU present")
    d = os.path.join(DATA, "18dof")
    s = SolverHalfImplicitGPU(18, 1)
    s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", 1, os.path.join(DATA, "input", "acc3.txt"))
    s.Initialize(s.zero_record(), 1)
    s.SetEndTime(1.0)
    with pytest.raises(LfdmError) as e:
        s.Solve()
    assert e.value.name == "LFDM_E_HIP"


@pytest.mark.parametrize("name", sorted(CASES1124[24]))
def test_initialize24_runs_tire_suspension_init(name, golden24):
    """Initialize on a 24-DOF handle applies initializeTireSus (dof24_halfImplicit_gpu.cu:320-333)
    on the host: zero records plus the case's initial speed give the reference's initial records bit
    for bit, with the per-vehicle masses of the knot cases (set before Initialize)."""
    from cases1124 import inputs

    flat, _, n, ovr, init = inputs(24, name, golden24)
    nr = bool(golden24[f"{name}/spec"][2])
    tt = TMeasyNr if nr else TMeasy
    s = SolverHalfImplicitGPU(24, n)
    s.construct_flat(flat, n, None, tt)
    if ovr is not None:
        s.set_overrides(ovr)
    rec = s.zero_record()
    rec[-24 + 3] = golden24[f"{name}/spec"][3]
    s.Initialize(rec, n)
    np.testing.assert_array_equal(s.get_states(), init)
    s.close()


def test_noalloc_refuses_fleets_beyond_its_array():
    """lfdm_integrate_noalloc checks the fleet against its fixed device array before any GPU work (no GPU
    is touched here: the pointers are never read)."""
    import ctypes

    from lfdm_amd._lib import lib

    flat = lfdm_amd.params_from_json(18, f"{DATA}/18dof/HMMWV/vehicle.json", f"{DATA}/18dof/HMMWV/tmeasy.json")
    rc = lib.lfdm_integrate_noalloc(18, flat.ctypes.data, flat.size, lfdm_amd.TMeasy, 200_000, 8, None, 8, 10, 1e-3,
                                    0, 8, ctypes.c_void_p(0))
    assert rc == -5
    assert "at most 64 MiB" in lib.lfdm_last_error().decode()
    rc = lib.lfdm_integrate_noalloc(18, flat.ctypes.data, flat.size, lfdm_amd.TMeasy, 0, 8, None, 8, 10, 1e-3, 0, 8,
                                    ctypes.c_void_p(0))
    assert rc == -1


def test_json_nesting_is_capped(tmp_path):
    """A parameter file nested beyond 64 levels is refused with a parse error, not a stack overflow."""
    deep = tmp_path / "deep.json"
    deep.write_text("[" * 100000 + "]" * 100000)
    veh = f"{DATA}/18dof/HMMWV/vehicle.json"
    with pytest.raises(LfdmError) as e:
        lfdm_amd.params_from_json(18, veh, str(deep))
    assert e.value.name == "LFDM_E_PARSE" and "nesting too deep" in str(e.value)


def test_parsing_ignores_the_process_locale():
    """Numbers parse the same under a comma-decimal C locale (run where such a locale is installed)."""
    import locale
    import subprocess
    import sys

    names = [n for n in ("de_DE.UTF-8", "de_DE.utf8", "fr_FR.UTF-8", "fr_FR.utf8", "de_DE", "fr_FR")]
    usable = None
    for n in names:
        try:
            old = locale.setlocale(locale.LC_NUMERIC)
            locale.setlocale(locale.LC_NUMERIC, n)
            locale.setlocale(locale.LC_NUMERIC, old)
            usable = n
            break
        except locale.Error:
            continue
    if usable is None:
        pytest.skip("no comma-decimal locale installed")
    veh, tire = f"{DATA}/18dof/HMMWV/vehicle.json", f"{DATA}/18dof/HMMWV/tmeasy.json"
    code = ("import locale, sys; sys.path.insert(0, %r); import lfdm_amd; locale.setlocale(locale.LC_ALL, %r); "
            "print(lfdm_amd.params_from_json(18, %r, %r).tobytes().hex())"
            % (os.path.join(ROOT, "low-fidelity-dynamic-models_amd"), usable, veh, tire))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True).stdout.strip()
    assert bytes.fromhex(out) == lfdm_amd.params_from_json(18, veh, tire).tobytes()


def test_vehicle_params_get_set_without_gpu():
    """lfdm_get_vehicle_params returns a vehicle's set with its per-vehicle parameters applied;
    lfdm_set_vehicle_params gives one vehicle any parameter vector (host side, before the device image:
    no GPU is touched)."""
    d = f"{DATA}/18dof"
    hm = lfdm_amd.params_from_json(18, f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json")
    s = SolverHalfImplicitGPU(18, 10)
    s.construct_flat(hm, 6, None, TMeasy)
    s.construct_flat(hm, 4, None, TMeasy)
    np.testing.assert_array_equal(s.get_vehicle_params(7), hm)
    from lfdm_amd import workloads as W

    ovr = W.default_overrides(hm, 10)
    ovr[0, 3] *= 1.1   # mass
    ovr[5, 3] = 0.5    # torque scale
    ovr[6, 3] = 777.0  # brake torque
    s.set_overrides(ovr)
    got = s.get_vehicle_params(3)
    F = W.FLAT18
    assert got[F["m"]] == hm[F["m"]] * 1.1 and got[F["maxBrakeTorque"]] == 777.0
    changed = np.flatnonzero(got != hm)
    assert set(changed) - {F["m"], F["maxBrakeTorque"]}  # the torque map ordinates, halved
    np.testing.assert_array_equal(s.get_vehicle_params(4), hm)
    mod = hm.copy()
    mod[F["h"]] += 0.1
    s.set_vehicle_params(2, mod)
    np.testing.assert_array_equal(s.get_vehicle_params(2), mod)  # its per-vehicle parameters follow the vector
    np.testing.assert_array_equal(s.get_vehicle_params(1), hm)
    s.set_vehicle_params(8, mod)  # joins vehicle 2's set (same vector, same driver table)
    np.testing.assert_array_equal(s.get_vehicle_params(8), mod)
    with pytest.raises(LfdmError):
        s.set_vehicle_params(10, mod)
    bad = mod.copy()
    bad[F["gears"]] = 0
    with pytest.raises(LfdmError):
        s.set_vehicle_params(0, bad)
    s.close()
