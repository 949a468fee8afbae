"""GPU: full-size instances of the batched workload, advanced together in one batched launch,
against the oracle's trajectories stored in tests/golden/c3_oracle.json (written by
tests/golden/make_c3_oracle.py: minutes per step on one CPU core for the largest of them)."""
import json
import os

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
    nsteps = min(len(gold[k]["steps"]) for k in keys)
    worst = 0.0
    for s in range(nsteps):
        batch.step(1)
        for k, ig in zip(keys, igs):
            ref = gold[k]["steps"][s]
            assert abs(ig.time() - ref["t"]) < 1e-12
            for op in "xyz":
                d = abs(ig.expect(0, SIGMA[op]).real - ref[op])
                worst = max(worst, d)
                assert d < 1e-10, (k, s, op, d)
            assert abs(ig.norm2() - ref["norm2"]) < 1e-10
    print(f"largest difference from the oracle: {worst:.2e}")
    batch.destroy()
