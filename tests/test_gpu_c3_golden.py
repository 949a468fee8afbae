"""GPU: full-size instances of the batched workload, advanced together in one batched launch,
against the oracle's trajectories stored in tests/golden/c3_oracle.json (written by
tests/golden/make_c3_oracle.py: seconds to minutes per step on one CPU core).  Each stored step
carries the oracle's own rounding spread (two runs from initial states perturbed by relative
1e-15); an instance must agree within max(1e-10, 10 x the largest spread so far), the bound the
metric's 1e-10 becomes for instances whose mostly unoccupied functions amplify rounding."""
import json
import os

import numpy as np
import pytest

from ttns_lib_amd.capi import SIGMA, Batch, Integrator, sb_problem
from ttns_lib_amd.workloads import GOLDEN, c3_instances

pytestmark = pytest.mark.gpu


def test_batched_instances_match_oracle_golden(hip):
    with open(os.path.join(GOLDEN, "c3_oracle.json")) as f:
        gold = json.load(f)
    insts = c3_instances()
    keys = sorted(gold, key=int)
    assert len(keys) >= 4
    keep, igs = [], []
    for k in keys:
        inst = insts[int(k)]
        tree, sop = sb_problem(hip, inst.model)
        ig = Integrator(hip, tree, sop, dt=inst.dt, krylov_dim=gold[k]["krylov_dim"],
                        krylov_tol=gold[k]["krylov_tol"])
        ig.set_sb_initial_state(inst.model)
        keep.append((tree, sop))
        igs.append(ig)
    batch = Batch(hip, igs)
    nsteps = max(len(gold[k]["steps"]) for k in keys)
    spread = {k: 0.0 for k in keys}
    worst = {k: 0.0 for k in keys}
    for s in range(nsteps):
        batch.step(1)
        for k, ig in zip(keys, igs):
            if s >= len(gold[k]["steps"]):
                continue
            ref = gold[k]["steps"][s]
            spread[k] = max(spread[k], ref.get("spread", 0.0))
            bound = max(1e-10, 10 * spread[k])
            assert abs(ig.time() - ref["t"]) < 1e-12
            for op in "xyz":
                d = abs(ig.expect(0, SIGMA[op]).real - ref[op])
                worst[k] = max(worst[k], d)
                assert d < bound, (k, s, op, d, bound)
            assert abs(ig.norm2() - ref["norm2"]) < 1e-10
    print("instance: steps, largest difference from the oracle, bound")
    for k in keys:
        print(f"  {k}: {len(gold[k]['steps'])}, {worst[k]:.2e}, {max(1e-10, 10 * spread[k]):.1e}")
    batch.destroy()
