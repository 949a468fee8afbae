"""GPU: the costate gradient of ttn_sb_adjoint_run against the independent-boson closed form,
against the oracle running the same method, against central differences of the trajectory
(every perturbed trajectory in one batched launch), and the autograd wrapper on device
tensors."""
import dataclasses
import time

import numpy as np
import pytest
import torch

from helpers import sb_model
from test_grad import DT, NSTEPS, G, W, closed_form, dephasing_model
from ttns_lib_amd.capi import SIGMA, Batch, Integrator, sb_problem
from ttns_lib_amd.grad import SigmaGradient, adjoint_gradient, observable_and_gradient
from ttns_lib_amd.workloads import c4_instance

pytestmark = pytest.mark.gpu


def test_gpu_adjoint_closed_form(hip):
    r = adjoint_gradient(hip, dephasing_model(), NSTEPS, DT, op="x", every=7)
    val, dg, dw = closed_form(W, G, NSTEPS * DT)
    assert abs(r.value - val) < 1e-12
    assert np.abs(r.dg - dg).max() < 1e-11 * np.abs(dg).max()
    assert np.abs(r.dw - dw).max() < 1e-11 * np.abs(dg).max()


def test_gpu_adjoint_matches_oracle(hip, orc):
    """Same method on both sides (6 modes, unoccupied functions, three checkpoint intervals):
    the trajectories agree to ~1e-13, the gradients to the rounding of ~10^3 quadrature terms."""
    m = sb_model(6, 6, 4, alpha=0.2)
    g = adjoint_gradient(hip, m, 20, 0.05, every=7)
    o = adjoint_gradient(orc, m, 20, 0.05, every=7)
    assert abs(g.value - o.value) < 1e-11
    scale = max(np.abs(o.dg).max(), np.abs(o.dw).max())
    assert np.abs(g.dg - o.dg).max() < 1e-9 * scale
    assert np.abs(g.dw - o.dw).max() < 1e-9 * scale


def selected_differences(lib, model, nsteps, dt, params, h):
    """central differences of <sigma_z(T)> in the parameters [(kind 'g'|'w', k)], all perturbed
    trajectories in one batched launch"""
    models = []
    for kind, k in params:
        for sgn in (1.0, -1.0):
            wk, gk = np.array(model.wk, float), np.array(model.gk, float)
            (gk if kind == "g" else wk)[k] += sgn * h
            models.append(dataclasses.replace(model, wk=wk, gk=gk))
    keep, igs = [], []
    for m in models:
        tree, sop = sb_problem(lib, m)
        ig = Integrator(lib, tree, sop, dt=dt, krylov_tol=1e-14)
        ig.set_sb_initial_state(m)
        keep.append((tree, sop))
        igs.append(ig)
    b = Batch(lib, igs)
    b.step(nsteps)
    b.destroy()
    vals = np.array([ig.expect(0, SIGMA["z"]).real for ig in igs])
    for ig in igs:
        ig.destroy()
    return (vals[0::2] - vals[1::2]) / (2 * h)


def test_gpu_adjoint_matches_batched_differences(hip):
    """The plain small configuration's tree (16 modes, boson 10, bond 8), 40 steps, two
    checkpoint intervals, against the batched central differences (h = 1e-4)."""
    m = sb_model(16, 10, 8)
    a = adjoint_gradient(hip, m, 40, 0.05, every=20)
    f = observable_and_gradient(hip, m, 40, 0.05, h=1e-4)
    av, fv = np.concatenate([a.dg, a.dw]), np.concatenate([f.dg, f.dw])
    assert abs(a.value - f.value) < 1e-12
    assert np.abs(av - fv).max() < 1e-5 * np.abs(fv).max()


def test_gpu_adjoint_on_the_adjoint_configuration(hip):
    """The 64-mode tree of the adjoint configuration (Ohmic alpha 0.2, wc 5, boson 12, bond 24),
    10 steps with checkpoints every 5 (two intervals), against central differences (h = 1e-4)
    in the couplings and frequencies of four modes across the bath."""
    inst = c4_instance()
    m = inst.model
    t0 = time.perf_counter()
    a = adjoint_gradient(hip, m, 10, inst.dt, every=5)
    t1 = time.perf_counter()
    print(f"\nc4 adjoint, 10 steps: {t1 - t0:.1f} s", flush=True)
    params = [("g", 0), ("g", 21), ("g", 42), ("g", 63), ("w", 0), ("w", 21), ("w", 42), ("w", 63)]
    fd = selected_differences(hip, m, 10, inst.dt, params, 1e-4)
    print(f"c4 central differences (16 trajectories, batched): {time.perf_counter() - t1:.1f} s", flush=True)
    ad = np.array([(a.dg if kind == "g" else a.dw)[k] for kind, k in params])
    rel = np.abs(ad - fd) / np.abs(fd).max()
    print(f"relative deviation {rel.max():.3e}\nadjoint {ad}\nfd      {fd}", flush=True)
    assert np.all(np.isfinite(a.dg)) and np.all(np.isfinite(a.dw))
    assert rel.max() < 1e-6  # measured 7.6e-8


def test_gpu_sigma_gradient_on_device_tensors(hip):
    g = torch.tensor(G, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(W, dtype=torch.float64, device="cuda", requires_grad=True)
    val, dg, dw = closed_form(W, G, NSTEPS * DT)
    out = SigmaGradient.apply(g, w, hip, dephasing_model(), NSTEPS, DT, "x")
    (2.0 * out).backward()
    assert out.device.type == "cuda" and g.grad.device.type == "cuda" and w.grad.device.type == "cuda"
    assert abs(out.item() - val) < 1e-12
    assert np.allclose(g.grad.cpu().numpy(), 2 * dg, rtol=0, atol=1e-10)
    assert np.allclose(w.grad.cpu().numpy(), 2 * dw, rtol=0, atol=1e-10)
