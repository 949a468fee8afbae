"""GPU: the HIP library against the reference's own outputs on its exact 512-mode runs (the
oracle passes the same checks in test_ref_outputs.py), and against the oracle."""
import pytest

from test_ref_outputs import (BOND_DIMENSION_RUNS, TRAJECTORY_CASES, check_trajectory, fixed_step_integrator,
                              ref_model, ref_runs)
from ttns_lib_amd.capi import SIGMA, sb_problem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", BOND_DIMENSION_RUNS)
def test_initial_bond_dimensions_match_reference_output(hip, name):
    run = ref_runs()[name]
    tree, _ = sb_problem(hip, ref_model(hip, run))
    n, _ = tree.size()
    assert [tree.node(i)["hrank"] for i in range(1, n)] == run["rows"][0]["hranks"]


# per-step bounds of the GPU against the oracle, ten to forty times the oracle's own spread under
# a 1e-15 relative perturbation of the initial state (subohmic: 1.4e-12, 6e-11; superohmic: 5e-9
# after one step; ohmic exponential: 2.8e-13, 3.6e-11, 4.6e-10; ohmic sudden: 1.5e-13, 2.5e-13,
# 1.0e-11): the GPU sums in another order everywhere, and any rebuild moves its rounding
ORACLE_BOUNDS = {"subohmic_exponential_t0": (2e-11, 5e-10), "superohmic_exponential": (5e-8,),
                 "ohmic_exponential": (1e-11, 1e-9, 1e-8), "ohmic_sudden": (1e-11, 1e-11, 3e-10)}


@pytest.mark.parametrize("name,steps,bound", TRAJECTORY_CASES)
def test_gpu_matches_reference_trajectory(hip, name, steps, bound):
    """Per step, the oracle's bound against the reference plus the GPU's bound against the oracle
    (the triangle inequality; e.g. the ohmic exponential run's third step: 1e-9 + 1e-8)."""
    print(f"[reference run] {name}, {steps} steps", flush=True)
    run = ref_runs()[name]
    _, _, ig = fixed_step_integrator(hip, run)
    check_trajectory(ig, run, steps, [bound + b for b in ORACLE_BOUNDS[name][:steps]])


@pytest.mark.parametrize("name", list(ORACLE_BOUNDS))
def test_gpu_matches_oracle_on_reference_run(hip, orc, name):
    run = ref_runs()[name]
    _, _, a = fixed_step_integrator(hip, run)
    _, _, o = fixed_step_integrator(orc, run)
    for bound in ORACLE_BOUNDS[name]:
        a.step(1)
        o.step(1)
        for op in "xyz":
            assert abs(a.expect(0, SIGMA[op]) - o.expect(0, SIGMA[op])) < bound
        assert abs(a.norm2() - o.norm2()) < 1e-12
