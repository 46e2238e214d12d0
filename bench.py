"""Throughput benchmark: 18-DOF FP64 vehicle-steps/s on 1,000,000 vehicles (one MI355X).

`--dof 11` / `--dof 24` run the same workload on the 11- and 24-DOF models (24-DOF vehicles start
at U(0, 20) m/s with rolling wheels); the default line is the 18-DOF headline.

Workload (knot_fleet): HMMWV vehicle + TMeasy tires, dt = 1e-3 s, mass / yaw inertia / cornering
stiffness jittered x U(0.9, 1.1) per vehicle, throttle piecewise constant U(0, 1) per 1 s,
steering U(-0.3, 0.3) rad knots every 0.5 s (linear), brake U(0, 0.5) on a seeded 10 % of the
1 s intervals.  One bench step integrates the whole fleet over a 10 s horizon (10,000 solver
steps) in one launch, with every input already resident in HBM.

Prints one JSON line; `roofline` uses the algorithmic FP64 FLOPs per vehicle-step counted by the
oracle's counting scalar on the same workload (+ - * / sqrt count 1, sin cos tan atan2 tanh
count 20, hypot counts 4) and the device time of the integration launch from HIP events on the
solver's stream.  `cpu_baseline` times the reference's own CPU solver (compiled from its sources
into oracle/_ref, one solver object per vehicle, Python threads) on a bounded sample of the same
workload, with the oracle (the CPU restatement, bit-identical to the reference solver, OpenMP over
vehicles) beside it under `port`; without oracle/_ref the oracle is the baseline.  The default
18-DOF line also carries `other_models` (the 24-DOF configuration, 500,000 vehicles over the same
horizon, and the 11-DOF model on the headline fleet), `adjoint` (the differentiation form on 65,536
vehicles x 5,000 steps, checkpoints every 100, with 20 MPC iterations; its headline
`adjoint_forward_cost_ratio` is (forward + backward) / plain forward at the named tape budget),
`mixed_fleet` (the batching form on 2,000,000 vehicles of the three models with horizons U(2, 20) s)
and `n_sweep` (the reference paper's fleet-size sweep on the acc3 driver file, with its A100 times).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "low-fidelity-dynamic-models_amd"))

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector / matrix peak
HBM_PEAK_GBS = 8000.0     # MI355X HBM3E


def flops_per_step(dof, flat, nsteps=10000, nveh=4, seed=2):
    """Algorithmic FLOPs per vehicle-step: + - * / sqrt count 1, sin/cos/tan/atan2/tanh count 20,
    hypot counts 4 (oracle counting scalar over the knot_fleet workload).  Returns (with every libm
    call the reference writes, without libm weights, with each step's distinct libm calls only: a call
    repeating one of the same step on the same argument counted once)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lfdm_oracle as O
    from lfdm_amd import workloads as W

    fl = W.knot_fleet(nveh, flat, seed=seed, dof=dof)
    init = initial_records(O, dof, flat, nveh, fl["ovr"])
    basic = trans = dtrans = 0
    for i in range(nveh):
        knots = fl["knots"][:, i:i + 1]
        c = O.count_ops(dof, flat, ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], knots), nsteps,
                        init=None if init is None else init[i])
        basic += c["add"] + c["mul"] + c["div"] + c["sqrt"]
        trans += 20 * (c["sin"] + c["cos"] + c["tan"] + c["atan2"] + c["tanh"]) + 4 * c["hypot"]
        dtrans += 20 * sum(c[f + "_distinct"] for f in ("sin", "cos", "tan", "atan2", "tanh")) + 4 * c["hypot_distinct"]
    steps = nsteps * nveh
    return (basic + trans) / steps, basic / steps, (basic + dtrans) / steps


def initial_records(O, dof, flat, n, ovr, seed=7):
    """24-DOF fleets start at U(0, 20) m/s with rolling wheels (workloads.initial_speed_records)
    after initializeTireSus; the other models start at rest (None)."""
    if dof != 24:
        return None
    from lfdm_amd import workloads as W

    return O.init24(flat, W.initial_speed_records(n, flat, dof, seed=seed), False, ovr)


def cpu_baseline(dof, flat, n_sample, nsteps, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import lfdm_oracle as O
    from lfdm_amd import workloads as W

    fl = W.knot_fleet(n_sample, flat, seed=2, dof=dof)
    ctrl = ("knots", fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    init = initial_records(O, dof, flat, n_sample, fl["ovr"])
    t0 = time.perf_counter()
    O.run(dof, flat, ctrl, n_sample, nsteps, ovr=fl["ovr"], init=init, threads=threads)
    dt = time.perf_counter() - t0
    # single-core rate on a small subset
    n1 = 64
    fl1 = W.knot_fleet(n1, flat, seed=3, dof=dof)
    init1 = initial_records(O, dof, flat, n1, fl1["ovr"])
    t1 = time.perf_counter()
    O.run(dof, flat, ("knots", fl1["ns"], fl1["ds"], fl1["np"], fl1["dp"], fl1["knots"]), n1, nsteps, ovr=fl1["ovr"],
          init=init1, threads=1)
    d1 = time.perf_counter() - t1
    return n_sample * nsteps / dt, n1 * nsteps / d1, dt


def cpu_reference(n_sample, nsteps, threads, n_single=32):
    """The reference's own 18-DOF CPU solver (d18SolverHalfImplicit::IntegrateStep, compiled from the
    reference sources into oracle/_ref/libref18.so by oracle/Makefile) on a sample of the knot_fleet
    workload: one solver object per vehicle with that vehicle's mass, yaw inertia and cornering
    stiffness, per-step controls from the vehicle's knot program.  Vehicles are spread over `threads`
    Python threads (ctypes releases the GIL during each call); construction and control expansion
    happen before the timed region.  Returns (vehicle-steps/s, single-core vehicle-steps/s, seconds)
    or None where oracle/_ref was not built."""
    import ctypes
    import threading

    from lfdm_amd import OVERRIDE_FIELDS
    from lfdm_amd import workloads as W

    path = os.path.join(ROOT, "oracle", "_ref", "libref18.so")
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    p, d, i, s = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_char_p
    L.ref18_create.restype = p
    L.ref18_create.argtypes = [s, s, i, s]
    L.ref18_destroy.argtypes = [p]
    L.ref18_set_step.argtypes = [p, d]
    L.ref18_set_param.restype = i
    L.ref18_set_param.argtypes = [p, s, d]
    L.ref18_run.restype = d
    L.ref18_run.argtypes = [p, d, i, p, i, p]
    hm = os.path.join(W.DATA, "18dof", "HMMWV")
    veh, tire = os.fsencode(f"{hm}/vehicle.json"), os.fsencode(f"{hm}/tmeasy.json")
    flat = hmmwv_params(18)
    t = W.time_grid(nsteps, 1e-3)

    def prepare(n, seed):
        fl = W.knot_fleet(n, flat, seed=seed, dof=18)
        hs, cs = [], []
        for v in range(n):
            h = L.ref18_create(veh, tire, 0, None)
            L.ref18_set_step(h, 1e-3)
            for f, name in enumerate(OVERRIDE_FIELDS):
                if L.ref18_set_param(h, name.encode(), float(fl["ovr"][f, v])) != 0:
                    raise RuntimeError(f"reference solver rejects parameter {name}")
            hs.append(h)
            cs.append(np.ascontiguousarray(
                W.knot_controls(t, fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"][:, v:v + 1])[:, :, 0]))
        return hs, cs

    def timed(hs, cs, nthreads):
        def work(k):
            for v in range(k, len(hs), nthreads):
                L.ref18_run(hs[v], 0.0, nsteps, cs[v].ctypes.data, 0, None)

        pool = [threading.Thread(target=work, args=(k,)) for k in range(nthreads)]
        t0 = time.perf_counter()
        for th in pool:
            th.start()
        for th in pool:
            th.join()
        dt = time.perf_counter() - t0
        for h in hs:
            L.ref18_destroy(h)
        return dt

    hs, cs = prepare(n_sample, 2)
    secs = timed(hs, cs, threads)
    del cs
    hs1, cs1 = prepare(n_single, 3)
    secs1 = timed(hs1, cs1, 1)
    return n_sample * nsteps / secs, n_single * nsteps / secs1, secs


def adjoint_cost(flat, n=65536, T=5000, K=100, mpc_iterations=20):
    """Differentiation form on the adjoint workload: 18-DOF HMMWV, n vehicles, T steps, tracking loss
    to a seeded arc at 15 m/s (lfdm_amd.mpc), checkpoints every K steps.  Device times of the
    adjoint's forward launch, of its backward launch and of the plain integration launch over the same
    controls, from HIP events on the solver's stream; then `mpc_iterations` Adam iterations of MPC."""
    import torch
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start, run_mpc
    from lfdm_amd.torch_ops import default_params, make_solver

    dev = "cuda"
    s = make_solver(flat, n)
    s.set_stream(torch.cuda.current_stream().cuda_stream)
    ref, kappa = arc_reference(n, T, device=dev)
    init = moving_start(flat, n)
    params = default_params(flat, n)
    s.set_overrides(params.numpy())
    s.set_states(init.numpy())
    ctrl = initial_controls(flat, n, T, kappa, device=dev)
    xy = torch.empty((T, 2, n), dtype=torch.float64, device=dev)
    gc = torch.empty_like(ctrl)
    gs = torch.empty((init.shape[1], n), dtype=torch.float64, device=dev)
    gp = torch.empty((7, n), dtype=torch.float64, device=dev)

    def pairs(budget):
        s.adjoint_set_tape_budget(budget)
        fwd, bwd = [], []
        tape = None
        for it in range(3):  # the first round warms up
            s.adjoint_forward(T, K, ctrl.data_ptr(), xy.data_ptr())
            f_ms = s.last_launch_ms()
            gxy = (2e-3 * (xy - ref)).contiguous()  # d/dxy of sum_k dt |xy_k - ref_k|^2
            tape = s.adjoint_tape_bytes()
            s.adjoint_backward(gxy.data_ptr(), gc.data_ptr(), gs.data_ptr(), gp.data_ptr())
            b_ms = s.last_launch_ms()
            if it:
                fwd.append(f_ms)
                bwd.append(b_ms)
        return float(np.mean(fwd)), float(np.mean(bwd)), tape

    # a 16 GiB tape budget (groups of 8 segments: the backward recomputes all but the last group's),
    # then the library's default, automatic budget (half of the free HBM: on MI355X it holds the whole
    # horizon's tape, so the forward tapes every step and the backward recomputes nothing)
    small_budget, nseg = 16 << 30, (T + K - 1) // K
    fd_ms, bd_ms, tape_small = pairs(small_budget)
    f_ms, b_ms, tape_auto = pairs(0)
    s.set_controls_steps_device(T, ctrl.data_ptr())
    plain = []
    for it in range(2):
        s.set_states(init.numpy())
        s.set_time(0.0)
        s.advance(T)
        plain.append(s.last_launch_ms())
    torch.cuda.synchronize()
    p_ms = float(np.min(plain))
    n_inputs = 3 * T + 6 + 23  # controls, parameters (m, Izz, Ixx, cornering stiffness, torque scale, brake), state
    if mpc_iterations:  # one untimed iteration: first-use allocations (gradients, optimizer state)
        run_mpc(s, flat, init, params, ref, ctrl, iterations=1)
        torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, hist = run_mpc(s, flat, init, params, ref, ctrl, iterations=mpc_iterations)
    torch.cuda.synchronize()
    mpc_s = time.perf_counter() - t0
    s.close()
    fd_acc = central_fd_accuracy(flat, T, K)
    auto_gb = tape_auto[0] / 1e9
    whole_horizon = tape_auto[1] >= nseg
    return {"vehicles": n, "steps": T, "checkpoint_every": K,
            "cost_ratio": (f_ms + b_ms) / p_ms,
            "cost_ratio_definition": f"(adjoint forward + backward) / plain forward (integration kernel on the same "
                                     f"controls), device times, at the library's default (automatic) tape budget: "
                                     f"{auto_gb:.1f} GB of tape, {tape_auto[1]} of {nseg} segments per backward group"
                                     + (" (the whole horizon: the forward tapes every step, the backward recomputes "
                                        "nothing)" if whole_horizon else ""),
            "tape_bytes": tape_auto[0], "segments_per_group": tape_auto[1], "forward_ms": f_ms, "backward_ms": b_ms,
            "plain_forward_ms": p_ms, "ratio_vs_taping_forward": (f_ms + b_ms) / f_ms,
            "budget_16GiB": {"tape_budget_bytes": small_budget, "tape_bytes": tape_small[0],
                             "segments_per_group": tape_small[1], "forward_ms": fd_ms, "backward_ms": bd_ms,
                             "cost_ratio": (fd_ms + bd_ms) / p_ms, "ratio_vs_taping_forward": (fd_ms + bd_ms) / fd_ms},
            "central_fd_cost_ratio": 2.0 * n_inputs * p_ms / (f_ms + b_ms),
            "central_fd_accuracy": fd_acc,
            "mpc": {"iterations": mpc_iterations, "seconds": mpc_s, "seconds_per_iteration": mpc_s / max(1, mpc_iterations),
                    "mean_loss_first": hist[0] if hist else None, "mean_loss_last": hist[-1] if hist else None},
            "note": "central_fd_cost_ratio = cost of central differences (2 plain forwards per input, "
                    f"{n_inputs} inputs per vehicle) / adjoint cost; central_fd_accuracy: the relative error central "
                    "differences reach at their best step against the adjoint (the adjoint itself matches the "
                    "dual-number oracle to ~1e-12, tests/test_gpu_adjoint.py)"}


def central_fd_accuracy(flat, T=5000, K=100, nb=4, seed=3):
    """Accuracy of central differences (J(x + h) - J(x - h)) / 2h of the adjoint workload's tracking loss,
    the gradient whose cost central_fd_cost_ratio prices: for nb vehicles of the workload (moving start,
    the MPC's feed-forward controls), 6 controls at random steps, the 7 per-vehicle parameters and 4
    initial states (u, v, wz, front-left wheel speed), each differenced at relative steps 1e-2 ... 1e-7
    on the GPU, against the adjoint's gradient.  Returns the best step's relative error per input
    (median and worst over inputs): what central differences can reach at all."""
    import torch

    import lfdm_amd
    from lfdm_amd import workloads as W
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start

    fields = lfdm_amd.record_fields(18)
    rng = np.random.default_rng(seed)
    ref, kappa = arc_reference(nb, T, device="cpu")
    ref = ref.numpy()
    c0 = initial_controls(flat, nb, T, kappa, device="cpu").numpy()
    p0 = W.default_overrides(flat, nb)
    i0 = moving_start(flat, nb).numpy()
    inputs = [("c", (int(j), int(q))) for j, q in zip(rng.integers(0, T, 6), rng.integers(0, 2, 6))]
    inputs += [("p", f) for f in range(7)] + [("s", fields.index(f)) for f in ("u", "v", "wz", "lf.omega")]
    hs = (1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-7)
    per = 1 + 2 * len(inputs) * len(hs)
    n = nb * per
    C = np.repeat(c0, per, axis=2)
    P = np.repeat(p0, per, axis=1)
    I = np.repeat(i0, per, axis=0)
    steps = {}
    for b in range(nb):
        for ii, (kind, d) in enumerate(inputs):
            x0 = c0[d[0], d[1], b] if kind == "c" else p0[d, b] if kind == "p" else i0[b, d]
            for hi, hr in enumerate(hs):
                h = hr * max(abs(x0), 1e-2 if kind != "p" else abs(x0))
                steps[(b, ii, hi)] = h
                for o, sign in enumerate((1.0, -1.0)):
                    col = b * per + 1 + 2 * (ii * len(hs) + hi) + o
                    if kind == "c":
                        C[d[0], d[1], col] += sign * h
                    elif kind == "p":
                        P[d, col] += sign * h
                    else:
                        I[col, d] += sign * h
    s = lfdm_amd.SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, lfdm_amd.TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_overrides(P)
    s.set_states(I)
    dc = torch.from_numpy(np.ascontiguousarray(C)).cuda()
    xy = torch.empty((T, 2, n), dtype=torch.float64, device="cuda")
    s.adjoint_forward(T, K, dc.data_ptr(), xy.data_ptr())
    s.synchronize()
    e = xy.cpu().numpy() - np.repeat(ref, per, axis=2)
    J = 1e-3 * (e * e).sum(axis=(0, 1))
    gxy = torch.from_numpy(np.ascontiguousarray(2e-3 * e)).cuda()
    gc = torch.empty_like(dc)
    gs = torch.empty((len(fields), n), dtype=torch.float64, device="cuda")
    gp = torch.empty((7, n), dtype=torch.float64, device="cuda")
    s.adjoint_backward(gxy.data_ptr(), gc.data_ptr(), gs.data_ptr(), gp.data_ptr())
    s.synchronize()
    s.close()
    gc, gs, gp = gc.cpu().numpy(), gs.cpu().numpy(), gp.cpu().numpy()
    best = []
    for b in range(nb):
        col0 = b * per
        for ii, (kind, d) in enumerate(inputs):
            adj = gc[d[0], d[1], col0] if kind == "c" else gp[d, col0] if kind == "p" else gs[d, col0]
            errs = []
            for hi in range(len(hs)):
                col = col0 + 1 + 2 * (ii * len(hs) + hi)
                fd = (J[col] - J[col + 1]) / (2 * steps[(b, ii, hi)])
                errs.append(abs(fd - adj) / max(abs(adj), 1e-300))
            best.append(min(errs))
    best = np.array(best)
    return {"inputs_per_vehicle": len(inputs), "vehicles": nb, "relative_steps": list(hs),
            "best_step_relative_error": {"median": float(np.median(best)), "max": float(best.max())}}


def n_sweep(H=10000, sizes=(1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072, 262144, 393216, 524288)):
    """The reference paper's N-sweep (paper/paper_results/bench.sh:32-42 with {11,18,24}dof_gpu_bench.cu):
    HMMWV with TMeasy tires on the acc3 driver file, Construct + Initialize at rest, 10 s at dt = 1e-3,
    output off, for each model and fleet size; one launch per run (device time from HIP events on the
    solver's stream, wall time around the launch), after a short warm-up launch from the same records.
    Beside it the reference's published A100 times of the same runs (paper/paper_results/data/LFDM),
    for context."""
    import lfdm_amd
    from lfdm_amd import workloads as W

    a100_ms = {  # paper/paper_results/data/LFDM/{11,18,24}dof.out, in the order of `sizes`
        11: (301.165, 244.829, 269.287, 320.415, 457.012, 1293.57, 2234.26, 4089.2, 8278.1, 12275, 16362.9),
        18: (485.186, 432.977, 460.563, 597.919, 982.565, 2821.58, 4877.27, 9150.85, 18626.2, 27482.3, 36857.3),
        24: (843.603, 794.223, 840.184, 1169.15, 2809.72, 6268.76, 13076.7, 24967.7, 49432.2, 73988, 98794.6)}
    out = {}
    acc3 = os.path.join(W.DATA, "input", "acc3.txt")
    for dof in (18, 11, 24):
        d = os.path.join(W.DATA, f"{dof}dof", "HMMWV")
        files = [f"{d}/vehicle.json", f"{d}/tmeasy.json"] + ([f"{d}/suspension.json"] if dof == 24 else [])
        rows = []
        for N, ref_ms in zip(sizes, a100_ms[dof]):
            s = lfdm_amd.SolverHalfImplicitGPU(dof, N)
            s.Construct(*files, N, acc3)
            s.Initialize(s.zero_record(), N)
            s.SetThreadsPerBlock(64)
            init = s.get_states()
            s.advance(200)  # warm-up: builds the device image
            s.set_states(init)
            s.set_time(0.0)
            s.synchronize()
            t0 = time.perf_counter()
            s.advance(H)
            s.synchronize()
            wall = time.perf_counter() - t0
            dev_ms = s.last_launch_ms()
            s.close()
            rows.append({"vehicles": N, "device_ms": dev_ms, "wall_ms": 1e3 * wall,
                         "vehicle_steps_per_s": N * H / wall, "a100_ms": ref_ms,
                         "a100_vehicle_steps_per_s": N * H / (ref_ms * 1e-3)})
        out[f"{dof}dof"] = rows
    return {"workload": "HMMWV TMeasy, acc3 driver file, start at rest, 10 s at dt=1e-3, output off "
                        "(the reference paper's GPU benchmark)", "sizes": list(sizes), "models": out}


def hmmwv_params(dof):
    import lfdm_amd
    from lfdm_amd import workloads as W

    d = os.path.join(W.DATA, f"{dof}dof", "HMMWV")
    return lfdm_amd.params_from_json(dof, f"{d}/vehicle.json", f"{d}/tmeasy.json",
                                     f"{d}/suspension.json" if dof == 24 else None)


def knot_fleet_rate(dof, flat, n, H, steps, warmup, tpb=64):
    """knot_fleet on n vehicles of the model, H solver steps per bench step (one launch), inputs
    resident in HBM: (vehicle-steps/s over the timed steps, their wall time, mean device time of a
    launch from HIP events on the solver's stream).  24-DOF vehicles start at U(0, 20) m/s."""
    from lfdm_amd import SolverHalfImplicitGPU, TMeasy
    from lfdm_amd import workloads as W

    fl = W.knot_fleet(n, flat, seed=2, dof=dof)
    s = SolverHalfImplicitGPU(dof, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.set_overrides(fl["ovr"])
    if dof == 24:
        s.initialize_batch(W.initial_speed_records(n, flat, dof, seed=7))
    else:
        s.Initialize(s.zero_record(), n)
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    s.SetThreadsPerBlock(tpb)
    del fl
    for _ in range(warmup):  # also uploads the fleet
        s.advance(H)
    s.synchronize()
    launch_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        s.advance(H)
        launch_ms.append(s.last_launch_ms())  # waits for this launch's end event
    s.synchronize()
    wall = time.perf_counter() - t0
    s.close()
    return n * H * steps / wall, wall, float(np.mean(launch_ms))


def mixed_fleet(n_total=2_000_000, split=(0.4, 0.4, 0.2), tmin=2.0, tmax=20.0, jitter=0.15, trial_steps=1000,
                reps=2, seed=11):
    """Batching form on the mixed workload: n_total vehicles split 40/40/20 over the 11-, 18- and
    24-DOF HMMWV models, per-vehicle horizons T_i ~ U(tmin, tmax) s (steps_i = round(T_i / dt)),
    mass / yaw inertia / cornering stiffness jittered x U(0.85, 1.15), knot-program controls over
    tmax; 11/18-DOF vehicles start at rest, 24-DOF ones at U(0, 20) m/s.  One bucket per model, the
    autotuner picks layout / block size / steps per launch, then each timed run is one graph
    launch over the longest horizon (records reset between runs, outside the timed region).
    Throughput counts only the steps each vehicle takes (sum of its horizon steps).  For
    comparison: the same fleet in vehicle order, one eager launch per model, each on its handle's
    own stream."""
    import lfdm_amd
    from lfdm_amd import LAYOUT_GROUPED, Fleet, SolverHalfImplicitGPU, TMeasy
    from lfdm_amd import workloads as W

    buckets, inits, horizons = [], [], []
    for m in W.mixed_fleet(n_total, split, tmin, tmax, jitter, seed):
        dof, n, flat, fl, hz = m["dof"], m["n"], m["flat"], m["fl"], m["horizon"]
        s = SolverHalfImplicitGPU(dof, n)
        s.construct_flat(flat, n, None, TMeasy)
        s.set_overrides(fl["ovr"])
        if dof == 24:
            s.initialize_batch(m["recs"])
        else:
            s.Initialize(s.zero_record(), n)
        s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
        del fl, m
        s.set_horizons(hz)
        buckets.append(s)
        inits.append(s.get_states())
        horizons.append(hz)
    T = int(max(h.max() for h in horizons))
    work = float(sum(h.astype(np.float64).sum() for h in horizons))

    def lane_efficiency(hz, sort):
        h = np.sort(hz)[::-1] if sort else hz
        pad = (-len(h)) % 64
        w = np.concatenate([h, np.zeros(pad, h.dtype)]).reshape(-1, 64).max(axis=1)
        return float(h.sum() / (64.0 * w.sum()))

    def reset():
        for s, i0 in zip(buckets, inits):
            s.set_states(i0)
            s.synchronize()

    f = Fleet(buckets)
    t0 = time.perf_counter()
    cfg = f.autotune(T, trial_steps=trial_steps)
    tune_s = time.perf_counter() - t0
    reset()
    f.run(T)  # captures the graph
    f.synchronize()
    ms, walls = [], []
    for _ in range(reps):
        reset()
        t0 = time.perf_counter()
        f.run(T)
        f.synchronize()
        walls.append(time.perf_counter() - t0)
        ms.append(f.last_ms())
    wall = min(walls)
    f.close()
    # comparison: the same buckets one after another, each a graph of its own with its autotuned
    # configuration and layout (what running the three models' fleets separately takes)
    seq = []
    for b, s in enumerate(buckets):
        g = Fleet([s])
        g.set_config(0, cfg[b][0], cfg[b][1], cfg[b][2])
        reset()
        g.run(T)  # captures the graph
        g.synchronize()
        reset()
        t0 = time.perf_counter()
        g.run(T)
        g.synchronize()
        seq.append(time.perf_counter() - t0)
        g.close()
    # comparison: vehicle order, one eager launch per model on each handle's own stream (the device
    # image in vehicle order is built, by a one-step launch, before the records are reset and timed)
    for s in buckets:
        s.set_layout(LAYOUT_GROUPED)
        s.set_time(0.0)
        s.advance(1)
    reset()
    t0 = time.perf_counter()
    for s in buckets:
        s.set_time(0.0)
        s.advance(T)
    for s in buckets:
        s.synchronize()
    eager_wall = time.perf_counter() - t0
    for s in buckets:
        s.close()
    return {"workload": f"mixed: {n_total:,} vehicles 11/18/24-DOF {split[0]:.0%}/{split[1]:.0%}/{split[2]:.0%}, "
                        f"horizons U({tmin:g}, {tmax:g}) s at dt=1e-3, +-{jitter:.0%} parameters, knot controls",
            "vehicle_steps": work, "longest_horizon_steps": T,
            "value": work / wall, "unit": "vehicle-steps/s", "graph_ms": float(min(ms)), "wall_s": wall,
            "autotune": {"seconds": tune_s, "trial_steps": trial_steps,
                         "buckets": [{"dof": d_, "vehicles": int(len(h)), "block": c[0], "steps_per_launch": c[1],
                                      "waves_per_simd": c[2], "layout": "by_horizon" if c[3] else "grouped"}
                                     for d_, h, c in zip((11, 18, 24), horizons, cfg)]},
            "lane_efficiency": {"by_horizon": [lane_efficiency(h, True) for h in horizons],
                                "grouped": [lane_efficiency(h, False) for h in horizons]},
            "buckets_one_after_another": {"wall_s": sum(seq), "per_bucket_s": seq,
                                          "graph_speedup": sum(seq) / wall},
            "grouped_eager": {"value": work / eager_wall, "wall_s": eager_wall},
            "speedup_vs_grouped_eager": eager_wall / wall}


def counters_figure(key, dof=18):
    """A figure of a model's integration-kernel counter runs summarized in profiles/ (None if the
    summary is absent): "traffic_bytes_per_vehicle_step" (calibrated FETCH_SIZE + WRITE_SIZE) or
    "valu_busy_per_simd" (SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES x waves per SIMD)."""
    path = os.path.join(ROOT, "profiles", f"integrate_{dof}dof_knot_fleet_counters.md")
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic_per_step():
    """HBM bytes per vehicle-step of the 18-DOF integration kernel (None if unmeasured)."""
    return counters_figure("traffic_bytes_per_vehicle_step")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--dof", type=int, default=18, choices=(11, 18, 24),
                    help="model; the headline line is 18-DOF (11/24 use the same generators, 24-DOF vehicles start "
                         "at U(0, 20) m/s)")
    ap.add_argument("--vehicles", type=int, default=1_000_000)
    ap.add_argument("--horizon-steps", type=int, default=10000)
    ap.add_argument("--tpb", type=int, default=64)
    ap.add_argument("--cpu-sample", type=int, default=32768)
    ap.add_argument("--ref-sample", type=int, default=4096,
                    help="vehicles of the reference CPU solver's timed sample (oracle/_ref, when built)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiation-form measurement")
    ap.add_argument("--adjoint-only", action="store_true", help="only the differentiation-form measurement")
    ap.add_argument("--mpc-iterations", type=int, default=20, help="MPC iterations of the adjoint measurement")
    ap.add_argument("--no-mixed", action="store_true", help="skip the batching-form (mixed fleet) measurement")
    ap.add_argument("--no-other-models", action="store_true", help="skip the 11- and 24-DOF lines of the default run")
    ap.add_argument("--mixed-only", action="store_true", help="only the batching-form measurement")
    ap.add_argument("--no-sweep", action="store_true", help="skip the reference paper's N-sweep (acc3, 1,024 .. "
                                                            "524,288 vehicles, all three models)")
    ap.add_argument("--sweep-only", action="store_true", help="only the N-sweep")
    ap.add_argument("--mixed-vehicles", type=int, default=2_000_000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (reads 1 GiB, writes 1 GiB), the "
                         "reference kernel that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU only; vehicles shard across GPUs with no "
                                   "communication between them, run with --gpus 1"}))
        sys.exit(2)

    if args.mixed_only:
        print(json.dumps(mixed_fleet(args.mixed_vehicles)))
        return

    if args.sweep_only:
        print(json.dumps(n_sweep(args.horizon_steps)))
        return

    if args.adjoint_only:
        print(json.dumps(adjoint_cost(hmmwv_params(18), mpc_iterations=args.mpc_iterations)))
        return

    n, H, dof = args.vehicles, args.horizon_steps, args.dof
    flat = hmmwv_params(dof)
    value, wall, kernel_ms = knot_fleet_rate(dof, flat, n, H, args.steps, args.warmup, args.tpb)
    ms_per_step = 1e3 * wall / args.steps
    fl_total, fl_basic, fl_exec = flops_per_step(dof, flat)
    achieved_tf = fl_total * n * H / (kernel_ms * 1e-3) / 1e12
    tps = measured_traffic_per_step() if dof == 18 else None
    roofline = {"bound": "fp64_valu", "achieved": achieved_tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved_tf / FP64_PEAK_TFLOPS, "traffic": None if tps is None else tps * n * H,
                "bound_note": "FP64 vector (VALU) issue: the step has no dense contraction and uses no MFMA; peak = "
                              "MI355X's published FP64 figure",
                "kernel": f"integrate<M{dof}, TMeasy, knots>", "kernel_ms": kernel_ms,
                "flop_per_vehicle_step": fl_total, "basic_flop_per_vehicle_step": fl_basic,
                "flop_per_vehicle_step_distinct_libm": fl_exec,
                "frac_basic_ops": fl_basic * n * H / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                "frac_distinct_libm": fl_exec * n * H / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                "traffic_source": "profiles/integrate_18dof_knot_fleet_counters.md (bytes per vehicle-step x "
                                  "vehicle-steps of one launch)" if tps is not None else None,
                "hbm_gbs": None if tps is None else tps * n * H / (kernel_ms * 1e-3) / 1e9,
                "hbm_frac": None if tps is None else tps * n * H / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "valu_busy_per_simd": counters_figure("valu_busy_per_simd") if dof == 18 else None,
                "valu_busy_source": "profiles/integrate_18dof_knot_fleet_counters.md (fraction of SIMD cycles "
                                    "issuing a VALU instruction; the step's bound)" if dof == 18 else None}
    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        rate, rate1, secs = cpu_baseline(dof, flat, args.cpu_sample, H, threads)
        port = {"value": rate, "unit": "vehicle-steps/s", "cores": threads, "kind": "port",
                "sample": f"knot_fleet, {args.cpu_sample} vehicles x {H} steps ({secs:.1f} s), "
                          f"oracle (bit-identical CPU restatement), OpenMP over vehicles",
                "single_core": rate1, "per_core": rate / threads, "cpu": cpu_model(), "nproc": os.cpu_count()}
        ref = cpu_reference(args.ref_sample, H, threads) if dof == 18 else None
        if ref is None:
            cpu = port
        else:
            cpu = {"value": ref[0], "unit": "vehicle-steps/s", "cores": threads, "kind": "reference",
                   "sample": f"knot_fleet, {args.ref_sample} vehicles x {H} steps ({ref[2]:.1f} s), the reference's "
                             f"own CPU solver (d18SolverHalfImplicit::IntegrateStep compiled from its sources into "
                             f"oracle/_ref), one solver object per vehicle, {threads} threads",
                   "single_core": ref[1], "per_core": ref[0] / threads, "cpu": cpu_model(),
                   "nproc": os.cpu_count(), "port": port}
    metric = ("18-DOF FP64 vehicle-steps/s at 1M vehicles; adjoint/forward cost ratio" if dof == 18 else
              f"{dof}-DOF FP64 vehicle-steps/s")
    other = None
    if dof == 18 and not args.no_other_models:
        # the 24-DOF configuration (500,000 vehicles, 10 s) and the 11-DOF model on the headline fleet
        other = {}
        for d_, n_ in ((24, 500_000), (11, 1_000_000)):
            v_, _, k_ = knot_fleet_rate(d_, hmmwv_params(d_), n_, H, 2, 1, args.tpb)
            f_, fb_, fe_ = flops_per_step(d_, hmmwv_params(d_))
            tf_ = lambda f: f * n_ * H / (k_ * 1e-3) / 1e12 / FP64_PEAK_TFLOPS  # noqa: E731
            other[f"{d_}dof"] = {"value": v_, "unit": "vehicle-steps/s", "vehicles": n_, "horizon_steps": H,
                                 "steps": 2, "kernel_ms": k_, "flop_per_vehicle_step": f_,
                                 "flop_per_vehicle_step_distinct_libm": fe_, "basic_flop_per_vehicle_step": fb_,
                                 "fp64_frac": tf_(f_), "fp64_frac_distinct_libm": tf_(fe_),
                                 "fp64_frac_basic_ops": tf_(fb_),
                                 "valu_busy_per_simd": counters_figure("valu_busy_per_simd", d_),
                                 "traffic_bytes_per_vehicle_step": counters_figure("traffic_bytes_per_vehicle_step", d_),
                                 "counters_source": (f"profiles/integrate_{d_}dof_knot_fleet_counters.md"
                                                     if counters_figure("valu_busy_per_simd", d_) is not None else None)}
    adj = None
    if dof == 18 and not args.no_adjoint:
        adj = adjoint_cost(flat, mpc_iterations=args.mpc_iterations)
    mixed = None
    if dof == 18 and not args.no_mixed:
        mixed = mixed_fleet(args.mixed_vehicles)
    sweep = None
    if dof == 18 and not args.no_sweep:
        sweep = n_sweep(H)
    line = {"metric": metric, "value": value,
            "unit": "vehicle-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"knot_fleet: {dof}-DOF HMMWV TMeasy, {n:,} vehicles, {H * 1e-3:g} s at dt=1e-3, "
                                   "+-10% mass/Izz/cornering stiffness, knot controls" +
                                   (", start at U(0,20) m/s" if dof == 24 else ""),
                       "vehicles": n, "horizon_steps": H, "dt": 1e-3, "threads_per_block": args.tpb},
            "roofline": roofline, "cpu_baseline": cpu, "adjoint": adj,
            "adjoint_forward_cost_ratio": None if adj is None else adj["cost_ratio"],
            "adjoint_cost_ratio_definition": None if adj is None else adj["cost_ratio_definition"],
            "mixed_fleet": mixed, "other_models": other, "n_sweep": sweep}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
