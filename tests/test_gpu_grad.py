"""GPU: central-difference gradients with every perturbed trajectory in one batched launch,
against the closed form and against the oracle's own central differences."""
import numpy as np
import pytest

from helpers import sb_model
from test_grad import DT, NSTEPS, check_closed_form, dephasing_model
from ttns_lib_amd.grad import observable_and_gradient

pytestmark = pytest.mark.gpu


def test_gpu_batched_gradient_closed_form(hip):
    r = observable_and_gradient(hip, dephasing_model(), NSTEPS, DT, op="x")
    assert r.batched and r.instances == 9
    check_closed_form(r)


def test_gpu_batched_gradient_matches_oracle(hip, orc):
    """Generic model (Delta = 1, 6 modes, unoccupied single-particle functions).  The central
    differences carry the trajectories' rounding noise / h; the bound is measured on the
    oracle itself as the change of its gradient between h = 1e-6 and h = 2e-6."""
    m = sb_model(6, 6, 4, alpha=0.2)
    g = observable_and_gradient(hip, m, 20, 0.05)
    o = observable_and_gradient(orc, m, 20, 0.05, batched=False)
    o2 = observable_and_gradient(orc, m, 20, 0.05, h=2e-6, batched=False)
    assert abs(g.value - o.value) < 1e-10
    spread = max(np.abs(o.dg - o2.dg).max(), np.abs(o.dw - o2.dw).max())
    scale = max(np.abs(o.dg).max(), np.abs(o.dw).max())
    assert spread < 1e-3 * scale
    assert np.abs(g.dg - o.dg).max() < 3 * spread + 1e-7
    assert np.abs(g.dw - o.dw).max() < 3 * spread + 1e-7
