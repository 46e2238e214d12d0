"""Batching form on the GPU: per-vehicle horizons, slot layouts, and mixed 11/18/24-DOF fleets run as
one HIP graph (lfdm_set_horizons, lfdm_set_layout, lfdm_fleet_*), through the C ABI.

Bars: a vehicle with horizon H ends bit-identical to the same vehicle run H steps without horizons
(and, for the reference's own acc3 trajectory, within 1e-9 normwise of the reference at step H);
layouts, graph replays, launch configurations and the autotuner never change a bit of any record.
"""
import os
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import LAYOUT_BY_HORIZON, LAYOUT_GROUPED, Fleet, SolverHalfImplicitGPU, TMeasy
from lfdm_amd import workloads as W
from cases18 import final_rel
from conftest import DATA

pytestmark = pytest.mark.gpu

TOL = 1e-9


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hmmwv_flat(dof):
    d = os.path.join(DATA, f"{dof}dof", "HMMWV")
    return lfdm_amd.params_from_json(dof, f"{d}/vehicle.json", f"{d}/tmeasy.json",
                                     f"{d}/suspension.json" if dof == 24 else None)


def build(dof, n, seed=4, horizons=None, layout=None, horizon_s=1.0):
    """HMMWV fleet of the knot_fleet workload (parameters jittered 15 %); 24-DOF starts moving."""
    flat = hmmwv_flat(dof)
    fl = W.knot_fleet(n, flat, seed=seed, horizon=horizon_s, jitter=0.15, dof=dof)
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.set_overrides(fl["ovr"])
    if dof == 24:
        s.initialize_batch(W.initial_speed_records(n, flat, dof, seed=seed))
    else:
        s.Initialize(s.zero_record(), n)
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    if layout is not None:
        s.set_layout(layout)
    if horizons is not None:
        s.set_horizons(horizons)
    return s


def truncated(dof, n, steps, seed=4):
    """Records of the fleet after `steps` steps without horizons."""
    r = build(dof, n, seed=seed)
    if steps > 0:
        r.advance(int(steps))
    out = r.get_states()
    r.close()
    return out


@pytest.mark.parametrize("layout", [LAYOUT_GROUPED, LAYOUT_BY_HORIZON])
@pytest.mark.parametrize("dof", [11, 18, 24])
def test_horizons_stop_each_vehicle_at_its_own_step(dof, layout):
    """200 vehicles (4 wavefronts, the last one padded), horizons from {0, 1, 37, 150, 299, 300, 5000}
    steps, 300 steps in two launches (120 + 180, so horizons straddle a launch boundary)."""
    say(f"[batching] horizons dof {dof} layout {layout}")
    n, T = 200, 300
    hz = np.random.default_rng(dof).choice([0, 1, 37, 150, 299, 300, 5000], n)
    s = build(dof, n, horizons=hz, layout=layout)
    assert s.get_layout() == layout
    s.advance(120)
    s.advance(180)
    got = s.get_states()
    s.close()
    for h in np.unique(hz):
        m = hz == h
        ref = truncated(dof, n, min(int(h), T))
        np.testing.assert_array_equal(got[m], ref[m], err_msg=f"horizon {h}")


def test_horizons_vs_reference_trajectory(golden18):
    """The reference's own HMMWV acc3 trajectory (sampled every 100 steps over 10 s): 100 vehicles,
    vehicle i stops after 100 (i + 1) steps and must equal the reference's state at that step."""
    say("[batching] horizons vs the reference's acc3 trajectory")
    flat = golden18["hmmwv_acc3/params"]
    traj = golden18["hmmwv_acc3/traj"][:, 0]  # (100, 75), sample k after 100 (k + 1) steps
    n = traj.shape[0]
    s = SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, lfdm_amd.load_driver_file(os.path.join(DATA, "input", "acc3.txt")), TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_horizons(100 * (np.arange(n) + 1))
    s.set_layout(LAYOUT_BY_HORIZON)
    s.SetEndTime(10.0)
    s.Solve()
    fin = s.get_states()
    s.close()
    err = final_rel(fin, traj, traj)
    say(f"[batching] acc3 horizons: max normwise rel {err.max():.3e}, bitwise {np.array_equal(fin, traj)}")
    assert err.max() <= TOL
    np.testing.assert_array_equal(fin[:, -1], traj[:, -1])  # gear


def read_csv(path):
    rows = open(path).read().splitlines()[1:]
    return np.array([[float(v) for v in r.rstrip(",").split(",")] for r in rows])


def test_output_rows_after_a_horizon_hold_the_final_state(tmp_path):
    """Solve with SetOutput at 100 Hz over 2 s in 0.5 s launches: rows up to a vehicle's horizon equal
    the run without horizons, later rows repeat the row at the horizon at their own times."""
    say("[batching] output rows with horizons")
    d = os.path.join(DATA, "18dof")
    n = 64
    hz = np.resize([0, 500, 770, 1250, 2000, 5000], n)
    for tag, h in (("ref", None), ("hz", hz)):
        s = SolverHalfImplicitGPU(18, n)
        s.Construct(f"{d}/HMMWV/vehicle.json", f"{d}/HMMWV/tmeasy.json", n, os.path.join(DATA, "input", "acc3.txt"))
        s.Initialize(s.zero_record(), n)
        s.SetEndTime(2.0)
        s.SetKernelSimTime(0.5)
        s.SetOutput(str(tmp_path / tag), 100, True, 0)
        if h is not None:
            s.set_horizons(h)
        s.Solve()
        s.close()
    for v in (0, 1, 2, 3, 4, 5, 33):
        a = read_csv(tmp_path / f"hz_{v}.csv")
        b = read_csv(tmp_path / f"ref_{v}.csv")
        assert a.shape == b.shape == (201, 13)
        np.testing.assert_array_equal(a[:, 0], b[:, 0])
        kh = min(int(hz[v]) // 10, 200)  # row index of the horizon (row k: after 10 k steps)
        np.testing.assert_array_equal(a[: kh + 1], b[: kh + 1])
        np.testing.assert_array_equal(a[kh + 1:, 1:], np.broadcast_to(b[kh, 1:], a[kh + 1:, 1:].shape))


MIX = ((11, 300, 11), (18, 300, 18), (24, 150, 24))


def mixed(with_horizons=True, T=400):
    out = []
    for dof, n, seed in MIX:
        hz = np.random.default_rng(seed).integers(0, 2 * T, n) if with_horizons else None
        out.append(build(dof, n, seed=seed, horizons=hz))
    return out


def test_fleet_graph_matches_eager_runs():
    """Mixed 11/18/24-DOF fleet with horizons: three graph replays (records reset between them, eager
    work in between) each bit-identical to eager launches; new launch configurations (block size,
    steps per launch, one to three waves per SIMD) recapture the graph and still give the same bits."""
    say("[batching] fleet graph vs eager")
    T = 400
    a, b = mixed(T=T), mixed(T=T)
    init = [s.get_states() for s in a]
    for s in b:
        s.advance(T)
    ref = [s.get_states() for s in b]
    f = Fleet(a)
    for rep in range(3):
        for s, i0 in zip(a, init):
            s.set_states(i0)
        f.run(T)
        f.synchronize()
        got = [s.get_states() for s in a]
        say(f"[batching] replay {rep}: {f.last_ms():.2f} ms")
        for (dof, _, _), g_, r_ in zip(MIX, got, ref):
            np.testing.assert_array_equal(g_, r_, err_msg=f"replay {rep}, {dof}-DOF")
    for c11 in ((256, 0, 1), (64, 100, 2), (128, 0, 3)):
        f.set_config(0, *c11)
        f.set_config(1, 64, 150, 2)
        f.set_config(2, 128, 0)
        for s, i0 in zip(a, init):
            s.set_states(i0)
        f.run(T)
        for (dof, _, _), s, r_ in zip(MIX, a, ref):
            np.testing.assert_array_equal(s.get_states(), r_, err_msg=f"reconfigured {c11}, {dof}-DOF")
    f.close()
    for s in a + b:
        s.close()


MAX_WAVES = {11: 4, 18: 3, 24: 2}


def test_large_fleet_graph_chains_buckets_and_matches_eager_runs():
    """Buckets that each fill the chip (on MI355X: at least 1,024 SIMDs x the kernel's waves per SIMD
    wavefronts) run one after another in the fleet's graph instead of side by side; with horizons,
    two replays, each bucket bit-identical to its own eager run."""
    T = 120
    sizes = ((11, 270_000, 5), (18, 200_000, 6), (24, 140_000, 7))
    say(f"[batching] large fleet graph ({', '.join(f'{n:,} {d}-DOF' for d, n, _ in sizes)}) vs eager, {T} steps")
    a, ref = [], []
    for dof, n, seed in sizes:
        hz = np.random.default_rng(seed).integers(T // 2, T + 1, n)
        a.append(build(dof, n, seed=seed, horizons=hz))
        r = build(dof, n, seed=seed, horizons=hz)
        r.advance(T)
        ref.append(r.get_states())
        r.close()
    init = [s.get_states() for s in a]
    f = Fleet(a)
    for rep in range(2):
        for s, i0 in zip(a, init):
            s.set_states(i0)
        f.run(T)
        f.synchronize()
        say(f"[batching] large fleet replay {rep}: {f.last_ms():.2f} ms")
        for (dof, _, _), s, r_ in zip(sizes, a, ref):
            np.testing.assert_array_equal(s.get_states(), r_, err_msg=f"replay {rep}, {dof}-DOF")
    f.close()
    for s in a:
        s.close()


@pytest.mark.parametrize("dof", [11, 18, 24])
def test_every_occupancy_gives_the_same_bits(dof):
    """The kernels built for one up to the model's most waves per SIMD run the same arithmetic: a fleet
    of 100,000 vehicles (two wavefronts per SIMD on MI355X, where the automatic choice is the kernel
    built for two) ends bit-identical under every occupancy and under the automatic choice."""
    n, T = 100_000, 300
    say(f"[batching] {dof}-DOF, {n} vehicles x {T} steps at 0 (automatic) .. {MAX_WAVES[dof]} waves per SIMD")
    s = build(dof, n)
    init = s.get_states()
    f = Fleet([s])
    ref = None
    for w in range(0, MAX_WAVES[dof] + 1):
        s.set_states(init)
        f.set_config(0, 64, 0, w)
        f.run(T)
        f.synchronize()
        got = s.get_states()
        if ref is None:
            ref = got
        else:
            np.testing.assert_array_equal(got, ref, err_msg=f"{dof}-DOF at {w} waves per SIMD")
    f.close()
    s.close()


def test_autotune_keeps_records_and_results():
    """The autotuner restores every record it trials on, picks the horizon layout for fleets with
    spread horizons, and the tuned graph gives the eager bits."""
    say("[batching] autotune")
    T = 400
    a, b = mixed(T=T), mixed(T=T)
    for s in b:
        s.advance(T)
    ref = [s.get_states() for s in b]
    before = [s.get_states() for s in a]
    f = Fleet(a)
    cfg = f.autotune(T, trial_steps=100)
    say(f"[batching] tuned (block, steps per launch, waves per SIMD, layout): {cfg}")
    for s, r_ in zip(a, before):
        np.testing.assert_array_equal(s.get_states(), r_)
    assert all(c[3] == LAYOUT_BY_HORIZON for c in cfg)
    assert all(c[0] in (64, 128, 256) for c in cfg)
    assert cfg[0][2] in (1, 2, 3) and cfg[1][2] in (1, 2) and cfg[2][2] == 1
    f.run(T)
    for (dof, _, _), s, r_ in zip(MIX, a, ref):
        np.testing.assert_array_equal(s.get_states(), r_, err_msg=f"{dof}-DOF")
    f.close()
    for s in a + b:
        s.close()
