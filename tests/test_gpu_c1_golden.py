"""GPU: the single larger run's tree (128 modes, boson dim 16, bond dim 32, dt 0.025; fixed-step
integrator, no subspace expansion) against the oracle's first steps stored in
tests/golden/c1_oracle.json (tests/golden/make_c1_oracle.py)."""
import json
import os

import pytest

from ttns_lib_amd.capi import SIGMA, Integrator, sb_problem
from ttns_lib_amd.workloads import GOLDEN, c1_instance

pytestmark = pytest.mark.gpu


def test_c1_tree_matches_oracle_golden(hip):
    with open(os.path.join(GOLDEN, "c1_oracle.json")) as f:
        gold = json.load(f)
    inst = c1_instance()
    assert abs(inst.dt - gold["dt"]) < 1e-15
    tree, sop = sb_problem(hip, inst.model)
    ig = Integrator(hip, tree, sop, dt=inst.dt, krylov_dim=gold["krylov_dim"], krylov_tol=gold["krylov_tol"])
    ig.set_sb_initial_state(inst.model)
    worst = 0.0
    for ref in gold["steps"]:
        ig.step(1)
        assert abs(ig.time() - ref["t"]) < 1e-12
        for op in "xyz":
            d = abs(ig.expect(0, SIGMA[op]).real - ref[op])
            worst = max(worst, d)
            assert d < 1e-10, (op, d)
        assert abs(ig.norm2() - ref["norm2"]) < 1e-10
    print(f"largest difference from the oracle: {worst:.2e}")
    ig.destroy()
