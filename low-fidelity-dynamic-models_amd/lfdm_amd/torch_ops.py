"""torch.autograd.Function over the 18-DOF differentiation form (HIP adjoint kernels).

``rollout(solver, controls, params, init, checkpoint_every)`` integrates a fleet of 18-DOF vehicles
from given initial records with per-step controls and per-vehicle parameters and returns the (x, y)
trajectory; its backward pass is the hand-written reverse-mode adjoint of the half-implicit step
(lfdm_adjoint_forward / lfdm_adjoint_backward, include/lfdm_amd.h), with checkpoints every
``checkpoint_every`` steps.  Gradients flow to controls, params and init.  Tensors are float64;
controls and the returned trajectory live on the GPU, params and init may live anywhere (they are
handed to the solver from device memory, so on the GPU nothing passes through the host).

The handle keeps one pending adjoint: a second rollout on the same solver (or new per-vehicle
parameters) before the first one's backward replaces it, and the first backward then raises instead of
returning gradients of the wrong trajectory.  Use one solver per rollout that is differentiated
together.
"""
import numpy as np
import torch

from .solver import OVERRIDE_FIELDS, SolverHalfImplicitGPU, TMeasy, record_len


class HalfImplicitRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, solver, controls, params, init, checkpoint_every):
        T, three, n = controls.shape
        if three != 3 or n != solver.total:
            raise ValueError(f"controls must be (steps, 3, {solver.total}), got {tuple(controls.shape)}")
        if controls.dtype != torch.float64 or not controls.is_cuda:
            raise ValueError("controls must be a float64 CUDA tensor")
        solver.set_stream(torch.cuda.current_stream(controls.device).cuda_stream)
        p_dev = params.detach().to(controls.device, torch.float64).contiguous()
        i_dev = init.detach().to(controls.device, torch.float64).contiguous()
        if tuple(p_dev.shape) != (len(OVERRIDE_FIELDS), n) or tuple(i_dev.shape) != (n, record_len(18, TMeasy)):
            raise ValueError("params must be (7, n) and init (n, record_len)")
        solver.set_overrides_device(p_dev.data_ptr())
        solver.set_states_device(i_dev.data_ptr())
        ctrl = controls.detach().contiguous()
        xy = torch.empty((T, 2, n), dtype=torch.float64, device=controls.device)
        solver.adjoint_forward(T, int(checkpoint_every), ctrl.data_ptr(), xy.data_ptr())
        ctx.generation = solver.adjoint_generation()
        ctx.keep = (p_dev, i_dev)  # read on the stream after this call returns
        ctx.solver = solver
        ctx.params_device = params.device
        ctx.init_device = init.device
        ctx.save_for_backward(ctrl)  # the kernels read it again in the backward pass
        return xy

    @staticmethod
    def backward(ctx, grad_xy):
        (ctrl,) = ctx.saved_tensors
        s = ctx.solver
        if s.adjoint_generation() != ctx.generation:
            raise RuntimeError("the solver ran another adjoint forward (or got new per-vehicle parameters) after "
                               "this rollout, so its backward state is gone; use one solver per rollout that is "
                               "differentiated together")
        g = grad_xy.contiguous().to(torch.float64)
        gctrl = torch.empty_like(ctrl)
        gstate = torch.empty((record_len(18, TMeasy), s.total), dtype=torch.float64, device=ctrl.device)
        gparam = torch.empty((len(OVERRIDE_FIELDS), s.total), dtype=torch.float64, device=ctrl.device)
        s.set_stream(torch.cuda.current_stream(ctrl.device).cuda_stream)
        s.adjoint_backward(g.data_ptr(), gctrl.data_ptr(), gstate.data_ptr(), gparam.data_ptr())
        return None, gctrl, gparam.to(ctx.params_device), gstate.t().to(ctx.init_device), None


def rollout(solver, controls, params, init, checkpoint_every=100):
    """(T, 2, n) positions after every step; differentiable w.r.t. controls (T, 3, n), params (7, n)
    and init (n, record_len)."""
    return HalfImplicitRollout.apply(solver, controls, params, init, checkpoint_every)


def make_solver(flat, n):
    """An 18-DOF TMeasy handle for n vehicles of one parameter set, ready for rollout()."""
    s = SolverHalfImplicitGPU(18, n)
    s.construct_flat(flat, n, None, TMeasy)
    s.Initialize(s.zero_record(), n)
    return s


def default_params(flat, n):
    from .workloads import default_overrides

    return torch.from_numpy(np.ascontiguousarray(default_overrides(flat, n)))
