"""Differentiation form on the GPU: the hand-written reverse-mode adjoint of the 18-DOF step
(csrc/d18_adjoint.h, kernels in csrc/d18_adjoint.hip) against fourth-order central finite
differences of the GPU forward map, through the C ABI and through torch.autograd.

The oracle pin: test_adjoint_matches_dual_oracle_on_tracking_workload compares the adjoint with the
oracle's forward-mode tangents (<= 1e-8 relative) on the tracking workload.  The finite-difference
check: for every input group (controls, per-vehicle parameters scaled by their values, initial state)
max |adjoint - fd| <= 1e-6 max |fd| (normwise relative error of the gradient); 1e-4 for the tire
deflections xe, ye, whose finite differences stop converging near 1e-5 (they are ~1e-5 m and stiff:
steps small enough to stay off the slip-force kinks come within a decade of the rounding noise;
the host check of the same code, tests/test_adjoint_host.py, reaches 8e-6).  FD uses (J(-2h) - 8 J(-h) + 8 J(h) - J(2h)) / 12h on a fixed random linear functional J
of the (x, y) trajectory, all perturbed copies of a vehicle integrated in one batch; steps: 1e-4 for
controls, 3e-6 relative for parameters, 1e-8 for deflections, 1e-7 for roll angle and rate, 1e-5
(relative above 1) for the other states.  The same check runs on the CPU over the kernels' own lane
code (tests/test_adjoint_host.py).
"""
import sys

import numpy as np
import pytest

import lfdm_amd
from lfdm_amd import workloads as W
from conftest import DATA

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def hmmwv():
    d = f"{DATA}/18dof/HMMWV"
    return lfdm_amd.params_from_json(18, f"{d}/vehicle.json", f"{d}/tmeasy.json")


def test_adjoint_matches_fourth_order_fd():
    from lfdm_amd.mpc import moving_start

    flat = hmmwv()
    fields = lfdm_amd.record_fields(18)
    rng = np.random.default_rng(3)
    nb, T, K = 3, 600, 64  # 10 segments, the last one partial
    dev = "cuda"
    # base vehicles: jittered parameters, moving starts, smooth random controls with braking
    base_p = W.default_overrides(flat, nb)
    base_p[0] *= rng.uniform(0.9, 1.1, nb)
    base_p[1] *= rng.uniform(0.9, 1.1, nb)
    base_p[3:5] *= rng.uniform(0.9, 1.1, nb)
    base_i = moving_start(flat, nb, speed=12.0).numpy()
    base_i[:, fields.index("v")] = 0.05
    base_i[:, fields.index("wz")] = 0.02
    t = np.arange(T) * 1e-3
    base_c = np.empty((T, 3, nb))
    for b in range(nb):
        base_c[:, 0, b] = 0.2 * np.sin(2.0 * t + b) + 0.05
        base_c[:, 1, b] = 0.35 + 0.15 * np.sin(3.0 * t + b)
        base_c[:, 2, b] = np.where(t > 0.4, 0.08, 0.0)
    w = rng.uniform(-1, 1, (T, 2, nb))
    # directions
    ctrl_dirs = [(int(j), int(q)) for j, q in zip(rng.integers(0, T, 8), rng.integers(0, 3, 8))] + [(T - 1, 0), (0, 1)]
    state_dirs = ["x", "y", "u", "v", "psi", "wz", "phi", "wx", "udot", "vdot", "crankOmega", "lf.omega", "rr.omega",
                  "lf.xe", "rf.ye", "lr.xe", "rr.ye"]
    dirs = [("c", d) for d in ctrl_dirs] + [("p", i) for i in range(7)] + [("s", f) for f in state_dirs]
    offs = (-2, -1, 1, 2)
    n = nb * (1 + 4 * len(dirs))
    P = np.repeat(base_p, 1 + 4 * len(dirs), axis=1)
    I = np.repeat(base_i, 1 + 4 * len(dirs), axis=0)
    C = np.repeat(base_c, 1 + 4 * len(dirs), axis=2)
    hs = []
    for b in range(nb):
        col0 = b * (1 + 4 * len(dirs))
        for di, (kind, d) in enumerate(dirs):
            if kind == "c":
                h = 1e-4
            elif kind == "p":
                h = 3e-6 * abs(base_p[d, b])
            else:
                x0 = base_i[b, fields.index(d)]
                h = (1e-8 if d.endswith((".xe", ".ye")) else 1e-7 if d in ("phi", "wx") else 1e-5) * max(1.0, abs(x0))
            hs.append(h)
            for oi, o in enumerate(offs):
                col = col0 + 1 + 4 * di + oi
                if kind == "c":
                    C[d[0], d[1], col] += o * h
                elif kind == "p":
                    P[d, col] += o * h
                else:
                    I[col, fields.index(d)] += o * h
    s = lfdm_amd.SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, lfdm_amd.TMeasy)
    s.set_overrides(P)
    s.initialize_batch(I)
    ctrl = torch.from_numpy(np.ascontiguousarray(C)).to(dev)
    xy = torch.empty((T, 2, n), dtype=torch.float64, device=dev)
    say(f"[adjoint] forward of {n} vehicles x {T} steps (checkpoint every {K})")
    s.adjoint_forward(T, K, ctrl.data_ptr(), xy.data_ptr())
    s.synchronize()
    W_all = np.repeat(w, 1 + 4 * len(dirs), axis=2)
    J = (xy.cpu().numpy() * W_all).sum(axis=(0, 1))
    xyb = torch.from_numpy(np.ascontiguousarray(W_all)).to(dev)
    gc = torch.empty_like(ctrl)
    gs = torch.empty((len(fields), n), dtype=torch.float64, device=dev)
    gp = torch.empty((7, n), dtype=torch.float64, device=dev)
    say("[adjoint] backward")
    s.adjoint_backward(xyb.data_ptr(), gc.data_ptr(), gs.data_ptr(), gp.data_ptr())
    s.synchronize()
    gc, gs, gp = gc.cpu().numpy(), gs.cpu().numpy(), gp.cpu().numpy()
    worst = {}
    k = 0
    for b in range(nb):
        col0 = b * (1 + 4 * len(dirs))
        err = {"c": [0.0, 0.0], "p": [0.0, 0.0], "s": [0.0, 0.0], "d": [0.0, 0.0]}
        for di, (kind, d) in enumerate(dirs):
            h = hs[k]
            k += 1
            Jd = J[col0 + 1 + 4 * di: col0 + 5 + 4 * di]
            fd = (Jd[0] - 8 * Jd[1] + 8 * Jd[2] - Jd[3]) / (12 * h)
            if kind == "c":
                ad, unit = gc[d[0], d[1], col0], 1.0
            elif kind == "p":
                ad, unit = gp[d, col0], abs(base_p[d, b])
            else:
                ad, unit = gs[fields.index(d), col0], 1.0
            g = "d" if kind == "s" and d.endswith((".xe", ".ye")) else kind
            err[g][0] = max(err[g][0], abs(ad - fd) * unit)
            err[g][1] = max(err[g][1], abs(fd) * unit)
        for kind, (num, den) in err.items():
            worst[kind] = max(worst.get(kind, 0.0), num / den)
    say(f"[adjoint] normwise relative error vs 4th-order FD: controls {worst['c']:.2e}, parameters "
        f"{worst['p']:.2e}, initial state {worst['s']:.2e}, tire deflections {worst['d']:.2e}")
    assert max(worst["c"], worst["p"], worst["s"]) <= 1e-6
    assert worst["d"] <= 1e-4
    # identical vehicles get identical gradients; untouched record fields get none
    assert np.all(gs[fields.index("lf.fx"), :] == 0.0)
    s.close()


def tracking_case(n, T, seed=21):
    """The differentiation-form workload at test size: HMMWV vehicles starting straight at 15 m/s,
    tracking a seeded arc (curvature U(-0.02, 0.02) 1/m at 15 m/s) with the loss of lfdm_amd.mpc; the
    per-vehicle parameters jittered x U(0.9, 1.1); controls the MPC's feed-forward guess plus smooth
    per-vehicle variations (steering, throttle, braking pulses), so gear, brake and slip states vary."""
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start

    flat = hmmwv()
    ref, kappa = arc_reference(n, T, seed=seed, device="cpu")
    rng = np.random.default_rng(seed)
    ovr = W.default_overrides(flat, n)
    ovr *= rng.uniform(0.9, 1.1, (7, 1)) * np.ones((1, n))
    ovr[:, :] *= rng.uniform(0.95, 1.05, (7, n))
    init = moving_start(flat, n).numpy()
    ctrl = initial_controls(flat, n, T, kappa, device="cpu").numpy()
    t = np.arange(T) * 1e-3
    ph = rng.uniform(0, 2 * np.pi, (3, n))
    ctrl[:, 0, :] += 0.05 * np.sin(1.3 * t[:, None] + ph[0])
    ctrl[:, 1, :] += 0.2 * np.sin(0.9 * t[:, None] + ph[1])
    ctrl[:, 2, :] = np.where(np.sin(1.7 * t[:, None] + ph[2]) > 0.8, 0.05, 0.0)
    return flat, ovr, init, np.ascontiguousarray(ctrl), ref.numpy()


def test_adjoint_matches_dual_oracle_on_tracking_workload():
    """The GPU adjoint's gradients of the arc tracking loss (5,000 steps, checkpoints every 100 steps,
    512 vehicles) against the oracle's forward-mode tangents (dual-number restatement of the reference
    step, oracle/orc_dual.h, pinned in tests/test_oracle_dual.py): per vehicle, the directional
    derivative along a random direction over the 15,000 controls, one over the 7 per-vehicle parameters
    and one over the initial record (tire deflections included), relative difference <= 1e-8.

    Both are exact derivatives of the piece of the discrete map the trajectory is on; the GPU's
    trajectory differs from the oracle's by device-libm rounding, so a vehicle whose derivative itself
    moves when a parameter moves by one ulp (the oracle's derivative spread; an event such as a gear
    shift lands on another step) is held to 25 x that spread instead, like the fleet parity tests."""
    import lfdm_oracle as O

    n, T, K, dt = 512, 5000, 100, 1e-3
    flat, ovr, init, ctrl, ref = tracking_case(n, T)
    rng = np.random.default_rng(8)
    fields = lfdm_amd.record_fields(18)
    dctrl = rng.standard_normal(ctrl.shape)
    dovr = rng.standard_normal(ovr.shape) * np.abs(ovr)
    scale = np.abs(init) + 1e-3
    scale[:, [i for i, f in enumerate(fields) if f.endswith((".xe", ".ye"))]] = 1e-5
    dinit = rng.standard_normal(init.shape) * scale
    say(f"[adjoint] tracking workload: {n} vehicles x {T} steps, K = {K}: GPU forward + backward")
    s = lfdm_amd.SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, lfdm_amd.TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_overrides(ovr)
    s.set_states(init)
    dc = torch.from_numpy(ctrl).cuda()
    xy = torch.empty((T, 2, n), dtype=torch.float64, device="cuda")
    s.adjoint_forward(T, K, dc.data_ptr(), xy.data_ptr())
    s.synchronize()
    e = xy.cpu().numpy() - ref
    loss_gpu = dt * (e * e).sum(axis=(0, 1))
    xyb = torch.from_numpy(np.ascontiguousarray(2.0 * dt * e)).cuda()
    gc = torch.empty_like(dc)
    gs = torch.empty((len(fields), n), dtype=torch.float64, device="cuda")
    gp = torch.empty((7, n), dtype=torch.float64, device="cuda")
    s.adjoint_backward(xyb.data_ptr(), gc.data_ptr(), gs.data_ptr(), gp.data_ptr())
    s.synchronize()
    s.close()
    gc, gs, gp = gc.cpu().numpy(), gs.cpu().numpy(), gp.cpu().numpy()
    gpu = np.stack([(gc * dctrl).sum(axis=(0, 1)), (gp * dovr).sum(axis=0), (gs.T * dinit).sum(axis=1)])
    say("[adjoint] oracle tangents and their one-ulp spread")
    loss, dl = O.loss_tangents18(flat, ovr, init, ctrl, ref, dctrl, dovr, dinit)
    spread = np.zeros((3, n))
    for f in (0, 1, 2, 3, 5, 6):
        for direction in (np.inf, -np.inf):
            o2 = ovr.copy()
            o2[f] = np.nextafter(o2[f], direction)
            _, d2 = O.loss_tangents18(flat, o2, init, ctrl, ref, dctrl, dovr, dinit)
            spread = np.maximum(spread, np.abs(d2 - dl) / np.abs(dl))
    assert np.isfinite(gpu).all() and np.isfinite(dl).all()
    rel = np.abs(gpu - dl) / np.abs(dl)
    stable = spread <= 1e-10
    bound = np.where(stable, 1e-8, np.maximum(1e-8, 25.0 * spread))
    lrel = np.abs(loss_gpu - loss) / loss
    for j, name in enumerate(("controls", "parameters", "initial record")):
        say(f"[adjoint] {name}: {stable[j].sum()} of {n} vehicles derivative-stable, max rel {rel[j][stable[j]].max():.2e}"
            f" (median {np.median(rel[j]):.1e}); others max rel / spread "
            f"{(rel[j][~stable[j]] / spread[j][~stable[j]]).max() if (~stable[j]).any() else 0:.2g}")
    say(f"[adjoint] loss: max rel difference {lrel.max():.2e}")
    assert (rel <= bound).all(), [np.flatnonzero(~(rel[j] <= bound[j]))[:10].tolist() for j in range(3)]
    assert stable.mean() > 0.9


def test_torch_rollout_backward_and_mpc_step():
    """rollout() is a torch.autograd.Function: loss.backward() fills controls.grad, params.grad and
    init.grad; gradients are linear in grad_output; a few MPC iterations reduce the tracking loss."""
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start, run_mpc, tracking_loss, \
        tracking_loss_and_grad
    from lfdm_amd.torch_ops import default_params, make_solver, rollout

    flat = hmmwv()
    n, T = 256, 1500
    s = make_solver(flat, n)
    ref, kappa = arc_reference(n, T, seed=4)
    init = moving_start(flat, n)
    params = default_params(flat, n)
    c0 = initial_controls(flat, n, T, kappa)
    ctrl = c0.clone().requires_grad_(True)
    p = params.clone().requires_grad_(True)
    i0 = init.clone().requires_grad_(True)
    xy = rollout(s, ctrl, p, i0, 100)
    loss = tracking_loss(xy, ref).sum()
    loss.backward()
    assert ctrl.grad is not None and p.grad is not None and i0.grad is not None
    assert torch.isfinite(ctrl.grad).all() and torch.isfinite(p.grad).all()
    g1 = ctrl.grad.clone()
    # scaling the loss scales the gradient exactly (the backward is linear in grad_output)
    ctrl.grad = None
    xy = rollout(s, ctrl, p, i0, 100)
    (2.0 * tracking_loss(xy, ref).sum()).backward()
    assert torch.equal(ctrl.grad, 2.0 * g1)
    # the MPC loop's loss gradient formed directly is autograd's, bit for bit
    ctrl.grad = None
    xy = rollout(s, ctrl, p, i0, 100)
    l2, g = tracking_loss_and_grad(xy, ref)
    xy.backward(g)
    assert torch.equal(ctrl.grad, g1)
    assert abs(l2.item() - loss.item()) <= 1e-12 * abs(loss.item())
    # checkpoint interval does not change the result
    ctrl.grad = None
    xy = rollout(s, ctrl, p, i0, 37)
    tracking_loss(xy, ref).sum().backward()
    assert torch.equal(ctrl.grad, g1)
    # nor does the backward's grouping: a tape budget below one segment makes every segment its
    # own recompute + reverse pair, the adjoints carried between them through device memory
    ctrl.grad = p.grad = i0.grad = None
    s.adjoint_set_tape_budget(1)
    xy = rollout(s, ctrl, p, i0, 100)
    assert s.adjoint_tape_bytes() == (100 * 38 * n * 8, 1)
    tracking_loss(xy, ref).sum().backward()
    assert torch.equal(ctrl.grad, g1)
    gp1, gi1 = p.grad.clone(), i0.grad.clone()
    # groups of 3 of the 15 segments (the forward tapes the last group, the backward recomputes the
    # other four), then the automatic default (half the free device memory), which holds all 15: the
    # forward tapes every step
    for budget, groups in ((3 * 100 * 38 * n * 8, 3), (0, 15)):
        ctrl.grad = p.grad = i0.grad = None
        s.adjoint_set_tape_budget(budget)
        xy = rollout(s, ctrl, p, i0, 100)
        assert s.adjoint_tape_bytes() == (groups * 100 * 38 * n * 8, groups)
        tracking_loss(xy, ref).sum().backward()
        assert torch.equal(ctrl.grad, g1) and torch.equal(p.grad, gp1) and torch.equal(i0.grad, gi1)
    say("[adjoint] MPC, 6 iterations")
    _, hist = run_mpc(s, flat, init, params, ref, c0, iterations=6, lr=0.05)
    say(f"[adjoint] MPC mean loss per vehicle: {[round(h, 4) for h in hist]}")
    assert hist[-1] < hist[0]
    s.close()


def test_two_rollouts_on_one_solver_refuse_a_stale_backward():
    """The handle keeps one pending adjoint: two rollouts on one solver summed into one loss make the
    first rollout's backward raise (its checkpoints were overwritten) instead of returning gradients of
    the wrong trajectory; so do new per-vehicle parameters between a forward and its backward.  With one
    solver per rollout the summed loss's gradients are the sum of the separate ones, bit for bit."""
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start, tracking_loss
    from lfdm_amd.torch_ops import default_params, make_solver, rollout

    flat = hmmwv()
    n, T = 128, 300
    ref, kappa = arc_reference(n, T, seed=6)
    init = moving_start(flat, n)
    params = default_params(flat, n)
    c0 = initial_controls(flat, n, T, kappa)
    s = make_solver(flat, n)
    ctrl = c0.clone().requires_grad_(True)
    loss = tracking_loss(rollout(s, ctrl, params, init), ref).sum() + \
        tracking_loss(rollout(s, 0.9 * ctrl, params, init), ref).sum()
    with pytest.raises(RuntimeError, match="another adjoint forward"):
        loss.backward()
    ctrl.grad = None
    loss = tracking_loss(rollout(s, ctrl, params, init), ref).sum()
    s.set_overrides(params.numpy())
    with pytest.raises(RuntimeError, match="another adjoint forward"):
        loss.backward()
    s2 = make_solver(flat, n)
    ctrl.grad = None
    (tracking_loss(rollout(s, ctrl, params, init), ref).sum() +
     tracking_loss(rollout(s2, 0.9 * ctrl, params, init), ref).sum()).backward()
    both = ctrl.grad.clone()
    ctrl.grad = None
    tracking_loss(rollout(s, ctrl, params, init), ref).sum().backward()
    g1 = ctrl.grad.clone()
    ctrl.grad = None
    tracking_loss(rollout(s2, 0.9 * ctrl, params, init), ref).sum().backward()
    assert torch.equal(both, g1 + ctrl.grad)
    s.close()
    s2.close()


def test_default_stream_orders_with_torch():
    """A handle handed torch's default stream (NULL) before its first launch launches on it: torch work
    queued after the adjoint forward sees the finished trajectory without an explicit synchronize,
    and the fill queued before it cannot land after it."""
    from lfdm_amd.mpc import arc_reference, initial_controls, moving_start
    from lfdm_amd.torch_ops import default_params, make_solver

    flat = hmmwv()
    n, T = 8192, 2000
    s = make_solver(flat, n)
    s.set_stream(torch.cuda.current_stream().cuda_stream)
    s.set_overrides(default_params(flat, n).numpy())
    s.set_states(moving_start(flat, n).numpy())
    _, kappa = arc_reference(n, T, seed=5)
    ctrl = initial_controls(flat, n, T, kappa)
    xy = torch.full((T, 2, n), float("nan"), dtype=torch.float64, device="cuda")
    s.adjoint_forward(T, 100, ctrl.data_ptr(), xy.data_ptr())
    early = xy.clone()  # queued on torch's stream right after the launch
    torch.cuda.synchronize()
    assert torch.isfinite(xy).all()
    assert torch.equal(early, xy)
    s.close()


@pytest.mark.parametrize("veh,tire", [("HMMWV/vehicle.json", "HMMWV/tmeasy.json"),
                                      ("audi/vehicle.json", "audi/tmeasy.json")])
def test_step_jacobian_matches_host_adjoint(veh, tire):
    """lfdm_step_jacobian (one step adjoint sweep per row on the GPU) against the same step adjoint
    compiled for the host (tests/cpp/jacobian_check.cpp, mode "asis"), vehicle by vehicle, at records
    reached by 3 s of knot-controlled driving and under per-vehicle random controls: row-wise
    normwise relative difference <= 1e-9 (host and device libm may differ in the last bit); the
    records do not change; without a torque converter (audi) the crankOmega column is zero.  The
    host step adjoint itself is checked against the reference's IntegrateStepWithJacobian and
    against finite differences (tests/test_adjoint_host.py)."""
    import os
    import subprocess

    from conftest import ROOT

    say(f"[adjoint] step Jacobian {veh}")
    d = f"{DATA}/18dof"
    flat = lfdm_amd.params_from_json(18, f"{d}/{veh}", f"{d}/{tire}")
    n = 300
    fl = W.knot_fleet(n, flat, seed=6, jitter=0.0)
    s = lfdm_amd.SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, lfdm_amd.TMeasy)
    s.Initialize(s.zero_record(), n)
    s.set_overrides(fl["ovr"])
    s.set_controls_knots(fl["ns"], fl["ds"], fl["np"], fl["dp"], fl["knots"])
    s.advance(3000)
    rec = s.get_states()
    rng = np.random.default_rng(11)
    ctrl = np.stack([rng.uniform(-0.4, 0.4, n), rng.uniform(0, 1, n), rng.uniform(0, 0.3, n)])
    dc = torch.tensor(ctrl, dtype=torch.float64, device="cuda")
    dj = torch.full((21, 24, n), float("nan"), dtype=torch.float64, device="cuda")
    rows, cols = s.step_jacobian(dc.data_ptr(), dj.data_ptr())
    s.synchronize()
    assert (rows, cols) == (21, 24)
    jac = dj.cpu().numpy()
    np.testing.assert_array_equal(s.get_states(), rec)
    s.close()
    assert np.isfinite(jac).all()
    tool = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "jacobian_check")
    worst = 0.0
    for v in (0, 37, 150, 299):
        inp = " ".join(repr(float(x)) for x in (*ctrl[:, v], 1e-3, *rec[v]))
        r = subprocess.run([tool, f"{d}/{veh}", f"{d}/{tire}", "asis"], input=inp, capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0, r.stderr
        host = np.array([[float(x) for x in line.split()] for line in r.stdout.splitlines()])
        assert host.shape == (21, 24)
        g = jac[:, :, v]
        err = np.abs(g - host).max(axis=1) / np.maximum(np.abs(host).max(axis=1), 1e-300)
        worst = max(worst, err.max())
    say(f"[adjoint] step Jacobian {veh}: worst row vs host {worst:.2e}")
    assert worst <= 1e-9
    if "audi" in veh:
        assert np.all(jac[:, 12, :] == 0.0)
