"""Gradients of a spin-boson TDVP observable with respect to the bath couplings g_k and
frequencies w_k.

adjoint_gradient: the costate method of ttn_sb_adjoint_run,

    d<O(T)>/dtheta = 2 int_0^T Im <chi(t)| dH/dtheta |psi(t)> dt,  chi(t) = exp(iH(T - t)) O psi(T),

one forward trajectory checkpointed every K steps, a second tree state chi propagated
backwards by the same TDVP step, and Gauss-Legendre quadrature of every step from the
leaves' transition densities; the cost does not grow with the number of parameters.

observable_and_gradient: central differences over 4 N + 1 trajectories (base, +-h in each
g_k and w_k) advanced in one batched launch (ttn_batch_step):

    d<O(T)>/dg_k ~ (<O(T)>[g_k + h] - <O(T)>[g_k - h]) / (2 h)

with O(h^2) truncation plus the trajectories' rounding noise divided by h, which is why the
Krylov tolerance defaults to 1e-14.  It checks the adjoint.

`SigmaGradient` wraps either as a torch.autograd.Function: forward returns <O(T)> for the
coupling and frequency tensors, backward returns the stored gradient times the incoming one.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .capi import SIGMA, Batch, Integrator, Library, SpinBoson, sb_problem


@dataclasses.dataclass
class GradientResult:
    value: float           # <O(T)> of the unperturbed model
    dg: np.ndarray         # d<O(T)>/dg_k
    dw: np.ndarray         # d<O(T)>/dw_k
    instances: int         # trajectories advanced
    batched: bool          # True when they ran in one batched launch


def _perturbed(model: SpinBoson, which: str, k: int, delta: float) -> SpinBoson:
    wk = np.array(model.wk, dtype=float)
    gk = np.array(model.gk, dtype=float)
    if which == "g":
        gk[k] += delta
    else:
        wk[k] += delta
    return dataclasses.replace(model, wk=wk, gk=gk)


def perturbed_models(model: SpinBoson, h: float, include_base: bool = True):
    """[model, g_0 + h, g_0 - h, ..., g_{N-1} - h, w_0 + h, ..., w_{N-1} - h] (without the
    leading unperturbed model when include_base is False)"""
    models = [model] if include_base else []
    for which in ("g", "w"):
        for k in range(len(model.wk)):
            models.append(_perturbed(model, which, k, +h))
            models.append(_perturbed(model, which, k, -h))
    return models


def observable_and_gradient(lib: Library, model: SpinBoson, nsteps: int, dt: float, *, op: str = "z",
                            h: float = 1e-6, krylov_dim: int = 12, krylov_tol: float = 1e-14,
                            batched: Optional[bool] = None, include_base: bool = True) -> GradientResult:
    """<sigma_op(T)> at T = nsteps * dt and its central-difference gradient in (g_k, w_k).
    batched=None batches whenever the library has the batch entry points.  With
    include_base=False the unperturbed trajectory is not run and the value is the mean of
    the g_0 +- h pair (an O(h^2) difference): 4 N trajectories, one per workgroup of the
    GPU for N = 64."""
    n = len(model.wk)
    models = perturbed_models(model, h, include_base)
    keep, igs = [], []
    for m in models:
        tree, sop = sb_problem(lib, m)
        ig = Integrator(lib, tree, sop, dt=dt, krylov_dim=krylov_dim, krylov_tol=krylov_tol)
        ig.set_sb_initial_state(m)
        keep.append((tree, sop))
        igs.append(ig)
    if batched is None:
        batched = lib.has("batch_create")
    try:
        if batched:
            b = Batch(lib, igs)
            try:
                b.step(nsteps)
            finally:
                b.destroy()
        else:
            for ig in igs:
                ig.step(nsteps)
        vals = np.array([ig.expect(0, SIGMA[op]).real for ig in igs])
    finally:
        for ig in igs:
            ig.destroy()
    if include_base:
        value, vals = float(vals[0]), vals[1:]
    else:
        value = float(0.5 * (vals[0] + vals[1]))
    d = (vals[0::2] - vals[1::2]) / (2.0 * h)
    return GradientResult(value, d[:n].copy(), d[n:].copy(), len(models), bool(batched))


def adjoint_gradient(lib: Library, model: SpinBoson, nsteps: int, dt: float, *, op: str = "z",
                     every: int = 20, nquad: int = 0, krylov_dim: int = 12,
                     krylov_tol: float = 1e-14) -> GradientResult:
    """<sigma_op(T)> at T = nsteps * dt and its gradient in (g_k, w_k) by the costate method
    (ttn_sb_adjoint_run): checkpoints every `every` steps, nquad-point quadrature per step
    (0: chosen from dt and the largest bath frequency)."""
    tree, sop = sb_problem(lib, model)
    ig = Integrator(lib, tree, sop, dt=dt, krylov_dim=krylov_dim, krylov_tol=krylov_tol)
    try:
        ig.set_sb_initial_state(model)
        value, dg, dw = ig.sb_adjoint(model, nsteps, 0, SIGMA[op], every=every, nquad=nquad)
    finally:
        ig.destroy()
    return GradientResult(float(value), dg, dw, 1, False)


try:
    import torch

    class SigmaGradient(torch.autograd.Function):
        """<sigma_op(T)> as a differentiable function of the couplings g and frequencies w
        (float64 tensors of the bath size); gradients by the costate method (method="adjoint",
        checkpoints every `every` steps) or batched central differences (method="fd", step h)."""

        @staticmethod
        def forward(ctx, g, w, lib, model, nsteps, dt, op="z", h=1e-6, krylov_tol=1e-14, method="adjoint",
                    every=20):
            m = dataclasses.replace(model, gk=g.detach().cpu().numpy().astype(float),
                                    wk=w.detach().cpu().numpy().astype(float))
            if method == "adjoint":
                r = adjoint_gradient(lib, m, nsteps, dt, op=op, every=every, krylov_tol=krylov_tol)
            elif method == "fd":
                r = observable_and_gradient(lib, m, nsteps, dt, op=op, h=h, krylov_tol=krylov_tol)
            else:
                raise ValueError(f"unknown gradient method {method!r}")
            ctx.save_for_backward(torch.as_tensor(r.dg, dtype=g.dtype, device=g.device),
                                  torch.as_tensor(r.dw, dtype=w.dtype, device=w.device))
            return torch.tensor(r.value, dtype=g.dtype, device=g.device)

        @staticmethod
        def backward(ctx, grad_out):
            dg, dw = ctx.saved_tensors
            return grad_out * dg, grad_out * dw, None, None, None, None, None, None, None, None, None
except ImportError:  # pragma: no cover - torch is part of the environment
    SigmaGradient = None
