"""GPU: single full-size instances against the oracle's first steps stored in
tests/golden/<name>_oracle.json (tests/golden/make_single_oracle.py): c1 = the single larger
run's tree (128 modes, boson dim 16, bond dim 32, dt 0.025; fixed-step integrator, no subspace
expansion), c4 = the gradient configuration's tree (64 modes, boson dim 12, bond dim 24)."""
import json
import os

import pytest

from ttns_lib_amd.capi import SIGMA, Integrator, sb_problem
from ttns_lib_amd.workloads import GOLDEN, c1_instance, c4_instance

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make", [("c1", c1_instance), ("c4", c4_instance)])
def test_single_instance_matches_oracle_golden(hip, name, make):
    with open(os.path.join(GOLDEN, f"{name}_oracle.json")) as f:
        gold = json.load(f)
    inst = make()
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
    print(f"{name}: largest difference from the oracle: {worst:.2e}")
    ig.destroy()
