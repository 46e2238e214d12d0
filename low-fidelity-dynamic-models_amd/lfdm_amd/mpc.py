"""Batched gradient-based MPC on the 18-DOF model: track a seeded reference path per vehicle by
gradient descent on the per-step controls, with gradients from the HIP adjoint (torch_ops.rollout).

Loss per vehicle (the integrand of the reference's FSA tracking cost, dof18_sundials.cpp:868-884):
    L = sum_k dt [(x_k - x_ref(t_k))^2 + (y_k - y_ref(t_k))^2],  t_k = (k + 1) dt,
with x_ref = sin(kappa v t) / kappa, y_ref = (1 - cos(kappa v t)) / kappa (a circular arc of
curvature kappa ~ U(-0.02, 0.02) 1/m driven at v = 15 m/s; a straight line as kappa -> 0).
"""
import numpy as np
import torch

from . import workloads as W
from .solver import record_fields
from .torch_ops import rollout

SPEED = 15.0


def arc_reference(n, nsteps, dt=1e-3, seed=11, speed=SPEED, device="cuda"):
    """(nsteps, 2, n) reference positions and the curvatures (n,)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = rng.uniform(-0.02, 0.02, n)
    t = (np.arange(nsteps) + 1.0) * dt
    s = speed * t[:, None]
    k = np.where(np.abs(kappa) < 1e-12, 1e-12, kappa)[None, :]
    ref = np.empty((nsteps, 2, n))
    ref[:, 0, :] = np.sin(k * s) / k
    ref[:, 1, :] = (1.0 - np.cos(k * s)) / k
    return torch.from_numpy(ref).to(device), kappa


def moving_start(flat, n, speed=SPEED, device="cpu"):
    """Initial records: straight ahead at `speed`, wheels rolling at speed / r0, first gear, crank
    5 % faster than the wheel speed over the first gear ratio (a slipping torque converter; with the
    crank exactly at that speed the converter's speed ratio is exactly 1, where the model switches
    between forward and reverse flow)."""
    fields = record_fields(18)
    rec = np.zeros((n, len(fields)))
    r0 = flat[W.tire_offsets(flat, False, 18)[0] - 25 + 3]
    om = speed / r0
    for i, f in enumerate(fields):
        if f.endswith(".omega"):
            rec[:, i] = om
    rec[:, fields.index("u")] = speed
    rec[:, fields.index("crankOmega")] = 1.05 * om / flat[W.FLAT18["gears"] + 1]
    return torch.from_numpy(rec).to(device)


def tracking_loss(xy, ref, dt=1e-3):
    """Per-vehicle tracking loss (n,)."""
    d = xy - ref
    return dt * (d * d).sum(dim=(0, 1))


def tracking_loss_and_grad(xy, ref, dt=1e-3):
    """tracking_loss summed over the vehicles (a scalar tensor) and its gradient with respect to xy,
    2 dt (xy - ref), in three passes over the trajectory (difference, dot product, in-place scaling);
    the gradient has the same bits as autograd's of tracking_loss(xy, ref).sum() (doubling is exact)."""
    with torch.no_grad():
        d = xy.detach() - ref
        v = d.view(-1)
        loss = dt * torch.dot(v, v)
        d.mul_(2.0 * dt)
    return loss, d


def initial_controls(flat, n, nsteps, kappa, device="cuda"):
    """Feed-forward guess: kinematic-bicycle steering for the arc's curvature, throttle 0.3."""
    wheelbase = flat[W.FLAT18["a"]] + flat[W.FLAT18["b"]]
    c = torch.zeros((nsteps, 3, n), dtype=torch.float64, device=device)
    c[:, 0, :] = torch.from_numpy(np.arctan(wheelbase * kappa) / flat[W.FLAT18["maxSteer"]]).to(device)[None, :]
    c[:, 1, :] = 0.3
    return c


def run_mpc(solver, flat, init, params, ref, ctrl0, iterations=20, lr=0.05, checkpoint_every=100, dt=1e-3):
    """Adam on the controls; returns (controls, per-iteration mean loss).  params and init move to the
    controls' device once; the iterations queue work without waiting for the GPU (the losses are read
    back after the last one).  The loss and its gradient with respect to the trajectory come from
    tracking_loss_and_grad (three passes over the (steps, 2, vehicles) trajectory instead of
    autograd's graph of the loss)."""
    ctrl = ctrl0.clone().requires_grad_(True)
    params = params.to(ctrl.device, torch.float64)
    init = init.to(ctrl.device, torch.float64)
    try:  # one fused kernel per step over the (steps, 3, vehicles) controls where torch has it
        opt = torch.optim.Adam([ctrl], lr=lr, fused=ctrl.is_cuda)
    except (RuntimeError, TypeError):
        opt = torch.optim.Adam([ctrl], lr=lr)
    # steering within [-1, 1], throttle and braking within [0, 1], in one pass
    lo = torch.tensor([-1.0, 0.0, 0.0], dtype=ctrl.dtype, device=ctrl.device).view(1, 3, 1)
    hi = torch.ones((1, 3, 1), dtype=ctrl.dtype, device=ctrl.device)
    history = []
    for _ in range(iterations):
        opt.zero_grad()
        xy = rollout(solver, ctrl, params, init, checkpoint_every)
        loss, g = tracking_loss_and_grad(xy, ref, dt)
        xy.backward(g)
        opt.step()
        with torch.no_grad():
            torch.clamp(ctrl, lo, hi, out=ctrl)
        history.append(loss.detach())
    losses = (torch.stack(history) / ctrl.shape[2]).tolist() if history else []
    return ctrl.detach(), losses
