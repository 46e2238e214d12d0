"""Pins the oracle's forward-mode dual scalar (oracle/orc_dual.h) and its tracking-loss tangents
(orc18_loss_tangents), which the GPU adjoint is checked against (test_gpu_adjoint.py).

- Every derivative rule at random points against the complex-step derivative Im f(x + ih) / h of the
  same function (exact to rounding for analytic functions), or sixth-order central differences where
  numpy has no complex form (atan2, hypot); fabs at 0 has derivative 0.
- The value part of the dual run is the double oracle's: the loss equals, bit for bit, the same sum
  over the double oracle's trajectory.
- The loss tangents along random directions over controls, per-vehicle parameters and the initial
  record match fourth-order central differences of the double oracle (the mathematics' derivative of
  the discrete map), to the accuracy of the differences.
"""
import cmath
import math

import numpy as np
import pytest

import lfdm_amd
import lfdm_oracle as O
from conftest import DATA
from lfdm_amd import workloads as W


def test_dual_rules_match_complex_step():
    rng = np.random.default_rng(1)
    h = 1e-30
    # math / cmath call the C library (glibc), as the oracle does
    for op, f, fc in ((0, math.sin, cmath.sin), (1, math.cos, cmath.cos), (2, math.tan, cmath.tan),
                      (3, math.tanh, cmath.tanh), (4, math.sqrt, cmath.sqrt)):
        for x in rng.uniform(0.05, 1.4, 50) * rng.choice([-1, 1], 50) if op != 4 else rng.uniform(1e-3, 50, 50):
            v, dx, _ = O.dual_op(op, float(x))
            assert v == f(x)
            cs = fc(complex(x, h)).imag / h
            assert abs(dx - cs) <= 1e-15 * max(1.0, abs(cs)), (op, x, dx, cs)
    for x, y in rng.uniform(-3, 3, (50, 2)).tolist():
        for op, f in ((8, lambda a, b: a / b), (9, lambda a, b: a * b)):
            v, dx, dy = O.dual_op(op, x, y)
            assert abs(dx - f(complex(x, h), y).imag / h) <= 1e-15 * max(1, abs(dx))
            assert abs(dy - f(x, complex(y, h)).imag / h) <= 1e-15 * max(1, abs(dy))
        for op, f in ((6, math.atan2), (7, math.hypot)):
            v, dx, dy = O.dual_op(op, x, y)
            assert v == f(x, y)
            e = 1e-3
            c = (-1, 9, -45, 0, 45, -9, 1)
            fdx = sum(ci * f(x + (k - 3) * e, y) for k, ci in enumerate(c)) / (60 * e)
            fdy = sum(ci * f(x, y + (k - 3) * e) for k, ci in enumerate(c)) / (60 * e)
            assert abs(dx - fdx) <= 1e-12 * max(1, abs(fdx)) and abs(dy - fdy) <= 1e-12 * max(1, abs(fdy))
    assert O.dual_op(5, 0.0)[1] == 0.0 and O.dual_op(5, -2.0)[1] == -1.0 and O.dual_op(5, 3.0)[1] == 1.0
    assert O.dual_op(7, 0.0, 0.0)[1:] == (0.0, 0.0)


def _case(n, T, seed=3):
    from lfdm_amd.mpc import moving_start

    d = f"{DATA}/18dof/HMMWV"
    flat = lfdm_amd.params_from_json(18, f"{d}/vehicle.json", f"{d}/tmeasy.json")
    rng = np.random.default_rng(seed)
    fields = lfdm_amd.record_fields(18)
    ovr = W.default_overrides(flat, n)
    ovr[0] *= rng.uniform(0.9, 1.1, n)
    ovr[3:5] *= rng.uniform(0.9, 1.1, n)
    init = moving_start(flat, n, speed=12.0).numpy()
    init[:, fields.index("v")] = 0.05
    t = np.arange(T) * 1e-3
    ctrl = np.empty((T, 3, n))
    for b in range(n):
        ctrl[:, 0, b] = 0.2 * np.sin(2.0 * t + b) + 0.05
        ctrl[:, 1, b] = 0.35 + 0.15 * np.sin(3.0 * t + b)
        ctrl[:, 2, b] = np.where(t > 0.2, 0.05, 0.0)
    ref = np.empty((T, 2, n))
    ref[:, 0, :] = 12.0 * (t[:, None] + 1e-3)
    ref[:, 1, :] = 0.1
    return flat, ovr, init, ctrl, ref


def _loss_double(flat, ovr, init, ctrl, ref, dt=1e-3):
    """The tracking loss from the double oracle's trajectory, summed in the dual run's order."""
    T, _, n = ctrl.shape
    _, tr = O.run(18, flat, ("steps", ctrl), n, T, ovr=ovr, init=init, every=1)
    fields = lfdm_amd.record_fields(18)
    x, y = tr[:, :, fields.index("x")], tr[:, :, fields.index("y")]
    L = np.zeros(n)
    for k in range(T):
        ex, ey = x[k] - ref[k, 0], y[k] - ref[k, 1]
        L = L + dt * (ex * ex + ey * ey)
    return L


def test_loss_tangents_vs_central_differences():
    n, T = 3, 400
    flat, ovr, init, ctrl, ref = _case(n, T)
    rng = np.random.default_rng(5)
    dctrl = rng.standard_normal(ctrl.shape)
    dovr = rng.standard_normal(ovr.shape) * np.abs(ovr) * 1e-2
    dinit = rng.standard_normal(init.shape) * (np.abs(init) + 1e-3)
    dinit[:, [i for i, f in enumerate(lfdm_amd.record_fields(18)) if f.endswith((".xe", ".ye"))]] *= 1e-4
    loss, dloss = O.loss_tangents18(flat, ovr, init, ctrl, ref, dctrl, dovr, dinit)
    np.testing.assert_array_equal(loss, _loss_double(flat, ovr, init, ctrl, ref))
    # fourth-order differences at a ladder of steps: each vehicle's best step (truncation against
    # rounding and the slip-force kinks) pins the tangent to ~1e-8
    steps = (1e-3, 3e-4, 1e-4, 3e-5, 1e-5, 3e-6, 1e-6)
    for j, pert in enumerate((lambda e: (ctrl + e * dctrl, ovr, init),
                              lambda e: (ctrl, ovr + e * dovr, init),
                              lambda e: (ctrl, ovr, init + e * dinit))):
        best = np.full(n, np.inf)
        for h in steps:
            vals = []
            for o_ in (-2, -1, 1, 2):
                c2, p2, i2 = pert(o_ * h)
                vals.append(_loss_double(flat, p2, i2, c2, ref))
            fd = (vals[0] - 8 * vals[1] + 8 * vals[2] - vals[3]) / (12 * h)
            best = np.minimum(best, np.abs(dloss[j] - fd) / np.abs(fd))
        print(f"direction {j}: dual {dloss[j]}, best relative difference to FD {best}")
        assert best.max() <= 1e-7, (j, best)


@pytest.mark.parametrize("seed", [1, 2])
def test_loss_tangents_are_linear_in_the_direction(seed):
    """dL along a + 2 b equals dL along a plus twice dL along b (to rounding), and the three directions
    are independent: seeding only one group leaves the others' tangents at zero."""
    n, T = 4, 300
    flat, ovr, init, ctrl, ref = _case(n, T, seed)
    rng = np.random.default_rng(seed)
    a = [rng.standard_normal(x.shape) for x in (ctrl, ovr, init)]
    b = [rng.standard_normal(x.shape) for x in (ctrl, ovr, init)]
    _, da = O.loss_tangents18(flat, ovr, init, ctrl, ref, *a)
    _, db = O.loss_tangents18(flat, ovr, init, ctrl, ref, *b)
    _, dab = O.loss_tangents18(flat, ovr, init, ctrl, ref, *[x + 2 * y for x, y in zip(a, b)])
    np.testing.assert_allclose(dab, da + 2 * db, rtol=1e-12, atol=1e-12 * np.abs(dab).max())
    z = [np.zeros_like(x) for x in a]
    _, d0 = O.loss_tangents18(flat, ovr, init, ctrl, ref, a[0], z[1], z[2])
    assert np.all(d0[1:] == 0) and np.array_equal(d0[0], da[0])
