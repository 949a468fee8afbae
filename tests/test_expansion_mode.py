"""The expansion integrator's mode (subspace_expansion_projector_splitting_integrator) on trees
whose bonds are all at capacity, where the reference's expansion returns without adding functions
and the step is the fixed-step one (bit for bit)."""
import numpy as np
import pytest

from helpers import make_integrator, sb_model
from ttns_lib_amd.capi import MODE_FIXED_STEP, MODE_SUBSPACE_EXPANSION, SIGMA


def _run(lib, model, mode, steps=6):
    ig = make_integrator(lib, model, mode=mode)[2]
    out = []
    for _ in range(steps):
        ig.step(1)
        out.append([ig.expect(0, SIGMA[o]) for o in "xyz"] + [ig.norm2()])
    nodes = [ig.get_node(i) for i in range(len(ig.shapes))]
    return np.array(out), nodes, ig.counters()[0]


def check_at_capacity_equals_fixed_step(lib):
    m = sb_model(8, 6, 4)
    a = _run(lib, m, MODE_FIXED_STEP)
    b = _run(lib, m, MODE_SUBSPACE_EXPANSION)
    assert np.array_equal(a[0], b[0])
    assert all(np.array_equal(x, y) for x, y in zip(a[1], b[1]))
    assert a[2] == b[2]


def test_oracle_at_capacity_equals_fixed_step(orc):
    check_at_capacity_equals_fixed_step(orc)


@pytest.mark.gpu
def test_gpu_at_capacity_equals_fixed_step(hip):
    check_at_capacity_equals_fixed_step(hip)
