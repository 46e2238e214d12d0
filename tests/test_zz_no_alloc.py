"""lfdm_integrate_noalloc (the entry point that allocates no device memory) against the handle path it
stands in for (construct_flat + initialize_batch + set_overrides + set_controls_steps_device + advance +
get_states), through the C ABI, bit for bit: two calls in a row on the same GPU, the second with fewer
vehicles (so it runs over what the first left in the library's device array), 18-DOF TMeasy in the fast
form and then 18-DOF in the strict form; then 11-DOF and 24-DOF TMeasyNr pairs.  (A fleet larger than
the array is refused before any GPU work: tests/test_capi.py.)"""
import ctypes
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import ARITH_FAST, ARITH_STRICT, SolverHalfImplicitGPU, TMeasy, TMeasyNr
from lfdm_amd import workloads as W
from lfdm_amd._lib import lib
from conftest import DATA

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def params(dof, nr):
    d = f"{DATA}/{dof}dof/" + ("Sedan" if nr else "HMMWV")
    tire = "tmeasyNr.json" if nr else "tmeasy.json"
    return lfdm_amd.params_from_json(dof, f"{d}/vehicle.json", f"{d}/{tire}", f"{d}/suspension.json" if dof == 24 else None,
                                     TMeasyNr if nr else TMeasy)


def inputs(dof, nr, n, nsteps, seed):
    """Records after a short knot-controlled drive (so every state field is nonzero), jittered
    parameters and random per-step controls."""
    flat = params(dof, nr)
    fl = W.knot_fleet(n, flat, seed=seed, dof=dof, tire_nr=nr)
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(flat, n, None, TMeasyNr if nr else TMeasy)
    if dof == 24:
        s.initialize_batch(W.initial_speed_records(n, flat, dof, seed=seed, tire_nr=nr))
    else:
        s.Initialize(s.zero_record(), n)
    s.set_overrides(fl["ovr"])
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    s.advance(700)
    recs = s.get_states()
    s.close()
    rng = np.random.default_rng(seed)
    ctrl = np.empty((nsteps, 3, n))
    ctrl[:, 0] = rng.uniform(-0.5, 0.5, (1, n)) * np.sin(np.linspace(0, 3, nsteps))[:, None]
    ctrl[:, 1] = rng.uniform(0.0, 1.0, (1, n))
    ctrl[:, 2] = np.where(rng.random((nsteps, n)) < 0.05, 0.2, 0.0)
    return flat, recs, fl["ovr"], ctrl


def handle_path(dof, nr, flat, recs, ovr, ctrl, arith):
    n = recs.shape[0]
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(flat, n, None, TMeasyNr if nr else TMeasy)
    s.set_states(recs)  # the records as they are (no 24-DOF initializeTireSus), like the no-alloc call
    s.set_overrides(ovr)
    s.set_arithmetic(arith)
    dc = torch.from_numpy(np.ascontiguousarray(ctrl)).cuda()
    s.set_controls_steps_device(ctrl.shape[0], dc.data_ptr())
    s.advance(ctrl.shape[0])
    out = s.get_states()
    s.close()
    return out


def noalloc(dof, nr, flat, recs, ovr, ctrl, arith):
    n = recs.shape[0]
    din = torch.from_numpy(np.ascontiguousarray(recs)).cuda()
    dov = torch.from_numpy(np.ascontiguousarray(ovr)).cuda()
    dc = torch.from_numpy(np.ascontiguousarray(ctrl)).cuda()
    dout = torch.full_like(din, float("nan"))
    f = np.ascontiguousarray(flat)
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.lfdm_integrate_noalloc(dof, f.ctypes.data, f.size, TMeasyNr if nr else TMeasy, n, din.data_ptr(),
                                    dov.data_ptr(), dc.data_ptr(), ctrl.shape[0], 1e-3, arith, dout.data_ptr(),
                                    ctypes.c_void_p(stream))
    assert rc == 0, lib.lfdm_last_error().decode()
    torch.cuda.synchronize()
    return dout.cpu().numpy()


@pytest.mark.parametrize("dof,nr,calls", [
    (18, False, ((3000, 1500, 1, ARITH_FAST), (700, 1000, 2, ARITH_STRICT))),
    (11, False, ((1000, 800, 3, ARITH_FAST), (130, 900, 4, ARITH_FAST))),
    (24, True, ((900, 600, 5, ARITH_STRICT), (64, 700, 6, ARITH_FAST))),
])
def test_noalloc_matches_handle_path(dof, nr, calls):
    for n, nsteps, seed, arith in calls:
        say(f"[noalloc] {dof}-DOF{' TMeasyNr' if nr else ''}: {n} vehicles x {nsteps} steps, arith {arith}")
        flat, recs, ovr, ctrl = inputs(dof, nr, n, nsteps, seed)
        want = handle_path(dof, nr, flat, recs, ovr, ctrl, arith)
        got = noalloc(dof, nr, flat, recs, ovr, ctrl, arith)
        np.testing.assert_array_equal(got, want)
        assert not np.array_equal(got, recs)
