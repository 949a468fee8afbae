"""Central-difference gradients of <sigma(T)> in (g_k, w_k) (ttns_lib_amd/grad.py) on the
oracle, pinned by the independent-boson closed form, and the torch.autograd wrapper."""
import numpy as np
import pytest
import torch

from ttns_lib_amd.capi import SpinBoson
from ttns_lib_amd.grad import SigmaGradient, observable_and_gradient

W = np.array([0.8, 2.0])
G = np.array([0.08, 0.15])
NSTEPS, DT = 40, 0.05


def dephasing_model(w=W, g=G):
    """Delta = 0, psi0 = (up + down)/sqrt(2): <sigma_x(t)> = prod_k exp(-4 g_k^2 (1 - cos w_k t) / w_k^2)
    exactly (no relative polaron phase), and two modes make the tree exact to rounding."""
    return SpinBoson(wk=np.array(w, float), gk=np.array(g, float), boson_dim=[12] * len(w), nspf=4, nspf_lower=4,
                     hsys=np.zeros((2, 2), complex), psi0=np.array([1, 1], complex) / np.sqrt(2))


def closed_form(w, g, t):
    f = np.exp(-4 * g * g * (1 - np.cos(w * t)) / (w * w))
    val = np.prod(f)
    dg = val * (-8 * g * (1 - np.cos(w * t)) / (w * w))
    dw = val * (-4 * g * g) * (t * np.sin(w * t) / (w * w) - 2 * (1 - np.cos(w * t)) / w ** 3)
    return val, dg, dw


def check_closed_form(r, value_tol=1e-12):
    val, dg, dw = closed_form(W, G, NSTEPS * DT)
    assert abs(r.value - val) < value_tol
    # FD truncation is O(h^2) ~ 1e-12; rounding noise / h ~ 1e-8
    assert np.all(np.abs(r.dg - dg) < 1e-7 * np.abs(dg).max())
    assert np.all(np.abs(r.dw - dw) < 1e-7 * np.abs(dg).max())
    # a swapped parameter or sign would be off by O(1)
    assert np.all(np.sign(r.dg) == np.sign(dg)) and np.all(np.sign(r.dw) == np.sign(dw))


def test_fd_gradient_closed_form(orc):
    r = observable_and_gradient(orc, dephasing_model(), NSTEPS, DT, op="x", batched=False)
    assert r.instances == 1 + 4 * len(W) and not r.batched
    check_closed_form(r)


def test_gradient_without_base_trajectory(orc):
    r = observable_and_gradient(orc, dephasing_model(), NSTEPS, DT, op="x", batched=False, include_base=False)
    assert r.instances == 4 * len(W)
    # the value is the mean of the g_0 +- h pair: off by h^2/2 d2<O>/dg_0^2 ~ 1e-11
    check_closed_form(r, value_tol=1e-10)


def test_autograd_function(orc):
    g = torch.tensor(G, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(W, dtype=torch.float64, requires_grad=True)
    val, dg, dw = closed_form(W, G, NSTEPS * DT)
    out = SigmaGradient.apply(g, w, orc, dephasing_model(), NSTEPS, DT, "x")
    (3.0 * out).backward()
    assert abs(out.item() - val) < 1e-12
    assert np.allclose(g.grad.numpy(), 3 * dg, rtol=0, atol=1e-6)
    assert np.allclose(w.grad.numpy(), 3 * dw, rtol=0, atol=1e-6)


def test_perturbation_targets_one_parameter(orc):
    from ttns_lib_amd.grad import _perturbed

    m = dephasing_model()
    p = _perturbed(m, "w", 1, 1e-3)
    assert np.array_equal(p.gk, m.gk) and p.wk[1] == m.wk[1] + 1e-3 and p.wk[0] == m.wk[0]
    p = _perturbed(m, "g", 0, -1e-3)
    assert np.array_equal(p.wk, m.wk) and p.gk[0] == m.gk[0] - 1e-3 and p.gk[1] == m.gk[1]
    # the caller's model is untouched
    assert np.array_equal(m.wk, W) and np.array_equal(m.gk, G)
