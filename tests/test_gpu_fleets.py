"""Fleet-scale parity on the GPU: every vehicle of seeded fleets of all three models and both tire
models over the full 10 s horizon (10,000 steps at dt = 1e-3), in both arithmetic forms, against the
oracle, under the per-vehicle bar of tests/fleetcheck.py (1e-9 for ulp-stable vehicles, 25 x the
model's own one-ulp response for the others, non-finite exactly where the oracle is wherever that is
ulp-stable).  Each test prints the number of vehicles outside 1e-9 and the bitwise fraction.

Fleets (the knot_fleet generator of the 10^6-vehicle throughput workload; 24-DOF vehicles start at
U(0, 20) m/s as in the 5x10^5-vehicle 24-DOF workload):
  18     18-DOF HMMWV TMeasy, 4,096 vehicles        18nr  18-DOF Sedan TMeasyNr, 2,048 vehicles
  11     11-DOF HMMWV TMeasy, 2,048 vehicles        11nr  11-DOF Sedan TMeasyNr, 2,048 vehicles
  24     24-DOF HMMWV TMeasy, 2,048 vehicles        24nr  24-DOF Sedan TMeasyNr, 2,048 vehicles
and the mixed fleet of the batching form (3,000 vehicles, 40/40/20 over 11/18/24-DOF HMMWV, parameters
x U(0.85, 1.15), horizons U(2, 20) s) run by Fleet.run as one autotuned graph launch, each vehicle
compared with the oracle truncated at its own horizon (the reference's multi-Construct fleets,
wheeled_vehicle_models/tests/18dof_gpu.cu:58-99, cannot express per-vehicle horizons).
"""
import functools
import os
import sys

import numpy as np
import pytest

import fleetcheck as F
from conftest import ROOT

pytestmark = pytest.mark.gpu

FLEETS = {  # tag: (dof, TMeasyNr?, vehicles, seed, golden case holding the parameters)
    "18": (18, False, 4096, 5, "hmmwv_acc3"),
    "18nr": (18, True, 2048, 9, "sedan_nr_acc05"),
    "11": (11, False, 2048, 5, "hmmwv_acc3"),
    "11nr": (11, True, 2048, 5, "sedan_nr_acc05"),
    "24": (24, False, 2048, 5, "hmmwv_acc3"),
    "24nr": (24, True, 2048, 5, "sedan_nr_acc05"),
}
NSTEPS, EVERY = 10000, 200


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@functools.lru_cache(maxsize=None)
def fleet(tag):
    import lfdm_oracle as O

    dof, nr, n, seed, case = FLEETS[tag]
    flat = np.load(os.path.join(ROOT, "tests", "golden", f"d{dof}_golden.npz"))[f"{case}/params"]
    c = F.build_fleet(dof, nr, n, flat, seed)
    say(f"[fleets] {tag}: oracle reference and one-ulp spread, {n} vehicles x {NSTEPS} steps")
    return c, F.reference(O, c, NSTEPS, EVERY)


@pytest.mark.parametrize("arith", ["fast", "strict"])
@pytest.mark.parametrize("tag", list(FLEETS))
def test_fleet_vs_oracle(tag, arith):
    case, R = fleet(tag)
    say(f"[fleets] {tag} {arith}: GPU run")
    gpu = F.gpu_run(case, NSTEPS, EVERY, F.ARITHS[arith])
    v = F.verdict(gpu, R)
    say("[fleets] " + F.summary_line(f"{tag} {arith}", v))
    assert v["ok"], v["failing"][:20]
    assert v["stable"] > 0.8 * v["n"]


@functools.lru_cache(maxsize=None)
def mixed():
    import lfdm_oracle as O

    models = F.mixed_inputs(3000)
    say("[fleets] mixed: oracle references at each vehicle's horizon")
    return models, [F.mixed_reference(O, m) for m in models]


@pytest.mark.parametrize("arith", ["fast", "strict"])
def test_mixed_fleet_graph_vs_oracle(arith):
    models, refs = mixed()
    say(f"[fleets] mixed {arith}: autotuned graph run")
    fins, cfg = F.mixed_gpu(models, F.ARITHS[arith])
    say(f"[fleets] mixed {arith}: tuned (block, steps per launch, waves per SIMD, layout) {cfg}")
    for m, fin, R in zip(models, fins, refs):
        v = F.final_verdict(fin, R)
        say("[fleets] " + F.summary_line(f"mixed {m['dof']}-DOF {arith}", v))
        assert v["ok"], (m["dof"], v["failing"][:20])
