"""The costate gradient (oracle/adjoint.cpp, the method of ttn_sb_adjoint_run) on the CPU:
transition densities against brute-force contraction of the full state vectors, the gradient
against the independent-boson closed form and against central differences of the trajectory,
checkpoint-interval independence and quadrature convergence."""
import dataclasses

import numpy as np
import pytest
import torch

from helpers import make_integrator, sb_model, tree_children, ttns_vector
from test_grad import DT, NSTEPS, G, W, closed_form, dephasing_model
from ttns_lib_amd.capi import SIGMA
from ttns_lib_amd.grad import SigmaGradient, adjoint_gradient, observable_and_gradient


def leaf_dims_preorder(tree):
    """(leaf modes in pre-order, their dimensions): the axis order of ttns_vector"""
    info, _ = tree_children(tree)
    modes = [x["leaf_index"] for x in info if x["nchild"] == 0]
    dims = [x["rows"] for x in info if x["nchild"] == 0]
    return modes, dims


def test_transition_densities_match_full_contraction(orc):
    """rho_m(j,l) = <bra|(|j><l|)_m|ket> and rhoS_m with Scoup on mode 0, for two different
    states (ket after 2 steps, bra after 3 steps backwards from another seed), from the full
    state vectors."""
    m = sb_model(3, 3, 3, alpha=0.3)
    tree, _, ket = make_integrator(orc, m)
    ket.step(2)
    m2 = dataclasses.replace(m, seed=7)
    tree2, _, bra = make_integrator(orc, m2)
    bra.step(3, coeff=1j)
    psi, chi = ttns_vector(ket, tree), ttns_vector(bra, tree2)
    modes, dims = leaf_dims_preorder(tree)
    P, X = psi.reshape(dims), chi.reshape(dims)
    S = np.asarray(m.scoup)
    out = ket.transition_densities(bra, 0, S)
    ax0 = modes.index(0)
    for mode, (rho, rhoS) in enumerate(out):
        ax = modes.index(mode)
        Pm, Xm = np.moveaxis(P, ax, 0), np.moveaxis(X, ax, 0)
        d = dims[ax]
        ref = Xm.conj().reshape(d, -1) @ Pm.reshape(d, -1).T
        assert np.abs(rho - ref).max() < 1e-13
        if mode == 0:
            assert np.all(rhoS == 0)
            continue
        SP = np.moveaxis(np.tensordot(S, P, axes=([1], [ax0])), 0, ax0)
        refS = np.moveaxis(SP, ax, 0).reshape(d, -1)
        refS = Xm.conj().reshape(d, -1) @ refS.T
        assert np.abs(rhoS - refS).max() < 1e-13
    # the trace of every leaf's density is the overlap
    ov = np.vdot(chi, psi)
    for rho, _ in out:
        assert abs(np.trace(rho) - ov) < 1e-13


def test_adjoint_closed_form(orc):
    """Independent-boson model (Delta = 0): the exact gradient of <sigma_x(T)>"""
    r = adjoint_gradient(orc, dephasing_model(), NSTEPS, DT, op="x", every=7)
    val, dg, dw = closed_form(W, G, NSTEPS * DT)
    assert abs(r.value - val) < 1e-12
    assert np.abs(r.dg - dg).max() < 1e-12 * np.abs(dg).max()
    assert np.abs(r.dw - dw).max() < 1e-12 * np.abs(dg).max()


@pytest.mark.parametrize("n,d,r,steps", [(2, 4, 4, 4), (6, 4, 4, 4), (8, 6, 6, 20)])
def test_adjoint_matches_central_differences(orc, n, d, r, steps):
    """Generic models (Delta = 1, unoccupied single-particle functions), against central
    differences of the trajectory itself (h = 1e-4: truncation ~1e-9, rounding ~1e-10).
    Measured: 2.7e-8 (complete two-mode tree), 1.6e-7 and 2.6e-6 relative (truncated trees,
    where the trajectory's derivative also carries the tree's projection error)."""
    m = sb_model(n, d, r)
    a = adjoint_gradient(orc, m, steps, 0.05, every=3)
    f = observable_and_gradient(orc, m, steps, 0.05, h=1e-4, batched=False)
    av, fv = np.concatenate([a.dg, a.dw]), np.concatenate([f.dg, f.dw])
    assert abs(a.value - f.value) < 1e-14
    bound = {2: 1e-7, 6: 1e-6, 8: 1e-5}[n]
    assert np.abs(av - fv).max() < bound * np.abs(fv).max()


def test_checkpoint_interval_does_not_change_the_gradient(orc):
    m = sb_model(4, 4, 4, alpha=0.3)
    rs = [adjoint_gradient(orc, m, 6, 0.05, every=k) for k in (1, 4, 6, 100)]
    for r in rs[1:]:
        assert r.value == rs[0].value
        assert np.array_equal(r.dg, rs[0].dg) and np.array_equal(r.dw, rs[0].dw)


def test_quadrature_converges(orc):
    """The step integrals converge with the Gauss rule.  A bath with 2 rad of phase per step
    (wc 5) is converged at 8 nodes; a stiff one (modes 20 and 40 on a complete tree) still
    moves between 4 and 16 nodes, and the automatic choice agrees with 16."""
    m = sb_model(4, 6, 4, alpha=0.5, wc=5.0)
    r8, r12 = (adjoint_gradient(orc, m, 5, 0.05, nquad=q) for q in (8, 12))
    v8, v12 = (np.concatenate([r.dg, r.dw]) for r in (r8, r12))
    assert np.abs(v8 - v12).max() < 1e-10 * np.abs(v12).max()
    stiff = sb_model(2, 6, 6, alpha=0.5, wc=20.0)
    r4, ra, r16 = (adjoint_gradient(orc, stiff, 5, 0.05, nquad=q) for q in (4, 0, 16))
    v4, va, v16 = (np.concatenate([r.dg, r.dw]) for r in (r4, ra, r16))
    assert np.abs(va - v16).max() < 1e-9 * np.abs(v16).max()
    assert np.abs(v4 - v16).max() > 1e-6 * np.abs(v16).max()


def test_adjoint_where_central_differences_fail(orc):
    """A stiff bath (modes 20-80) on a truncated tree: the trajectory's central differences do
    not converge (h = 1e-3 ... 1e-6 give O(1) different answers: mostly unoccupied
    single-particle functions make the step's gauge rounding-chaotic), while the costate
    gradient is smooth in the quadrature and within 30 % of the h = 1e-3 difference (measured 26 %)."""
    m = sb_model(4, 6, 4, alpha=0.5, wc=20.0)
    a = adjoint_gradient(orc, m, 5, 0.05)
    fds = [observable_and_gradient(orc, m, 5, 0.05, h=h, batched=False) for h in (1e-3, 1e-5)]
    assert np.abs(fds[0].dg - fds[1].dg).max() > 0.1 * np.abs(a.dg).max()
    assert np.abs(a.dg - fds[0].dg).max() < 0.35 * np.abs(a.dg).max()


def test_adjoint_refusals(orc):
    m = sb_model(2, 4, 4)
    _, _, ig = make_integrator(orc, m)
    with pytest.raises(Exception):
        ig.sb_adjoint(m, 0, 0, SIGMA["z"])
    with pytest.raises(Exception):
        ig.sb_adjoint(m, 2, 0, SIGMA["z"], nquad=17)
    with pytest.raises(Exception):
        ig.sb_adjoint(m, 2, 0, SIGMA["z"], every=0)
    ig.step(1)
    with pytest.raises(Exception, match="stepped"):
        ig.sb_adjoint(m, 2, 0, SIGMA["z"])
    other = sb_model(3, 4, 4)
    _, _, ig2 = make_integrator(orc, m)
    with pytest.raises(Exception):
        ig2.sb_adjoint(other, 2, 0, SIGMA["z"])


def test_autograd_fd_and_adjoint_agree(orc):
    g = torch.tensor(G, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(W, dtype=torch.float64, requires_grad=True)
    out = SigmaGradient.apply(g, w, orc, dephasing_model(), NSTEPS, DT, "x", 1e-6, 1e-14, "fd")
    out.backward()
    ga, wa = g.grad.clone(), w.grad.clone()
    g.grad = None
    w.grad = None
    out2 = SigmaGradient.apply(g, w, orc, dephasing_model(), NSTEPS, DT, "x", 1e-6, 1e-14, "adjoint", 5)
    out2.backward()
    assert abs(out.item() - out2.item()) < 1e-12
    assert torch.allclose(ga, g.grad, rtol=0, atol=1e-7) and torch.allclose(wa, w.grad, rtol=0, atol=1e-7)
    with pytest.raises(ValueError):
        SigmaGradient.apply(g, w, orc, dephasing_model(), 2, DT, "x", 1e-6, 1e-14, "nope")
