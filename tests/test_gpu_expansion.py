"""GPU: the subspace expansion of the expansion integrator inside the persistent sweep, against
the oracle and against mathematics.

* full-rank limit: from bond dimension 2, or 1 with the random one-site expansion, the bonds grow
  to the Schmidt ranks and the trajectory equals the dense propagator (1e-12);
* the reference's own runs: after the steps in which the oracle reproduces the reference's printed
  counters and all 1024 bond dimensions exactly, the HIP path does too, and the Debye run's growth
  stays within the oracle's bounds around the reference's;
* against the oracle on the same inputs and the same normal stream, step by step, bounded by how
  far the oracle drifts from itself under 1e-15 relative perturbations of the initial node
  tensors (helpers.RoundingYardstick): identical bond dimensions and counters while every
  spawning decision so far was robust to that rounding and every perturbed run kept them, the
  drift of the bonds afterwards within twice the perturbed runs' plus a slack, and <sigma> within
  max(1e-10, 10 x the perturbed spread); small trees on one workgroup and on a team, and the
  single larger run (128 modes, bond dimension 8 grown to capacity 32 over 50 steps) against
  tests/golden/c1x_expansion.npz.  The two-site spectra of bonds whose newest functions carry
  almost no weight move by factors of 10 and more under such perturbations (step 5 onwards in
  the single larger run), so only robust steps can be asked to match exactly;
* a limit that adds nothing (infinite spawning threshold) equals the fixed-step integrator bit
  for bit, and a tree at capacity likewise (tests/test_expansion_mode.py).
"""
import os

import numpy as np
import pytest

from helpers import RoundingYardstick, exact_sigma, make_integrator, sb_model, ttns_vector
from test_expansion import EXACT_GROWTH, _full_rank_model, _state, check_debye_growth, \
    check_no_spawn_equals_fixed_step, check_reference_growth, schmidt_ranks, side_by_side
from ttns_lib_amd.capi import MODE_SUBSPACE_EXPANSION, SIGMA, Integrator, sb_problem
from ttns_lib_amd.workloads import C1_EXPANSION, GOLDEN, c1_expansion_instance

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nmodes,r0,min_unocc", [(2, 2, 0), (3, 2, 0), (4, 2, 0), (4, 1, 2)])
def test_gpu_expansion_reaches_full_rank_and_the_dense_propagator(hip, nmodes, r0, min_unocc):
    print(f"[gpu expansion full rank] modes {nmodes} r0 {r0} min_unocc {min_unocc}", flush=True)
    model = _full_rank_model(nmodes, r0)
    tree, _, ig = make_integrator(hip, model, dt=0.05, krylov_dim=16, krylov_tol=1e-13, mode=MODE_SUBSPACE_EXPANSION,
                                  spawning_threshold=-1.0, unoccupied_threshold=1e-4, minimum_unoccupied=min_unocc)
    psi0 = ttns_vector(ig, tree)
    times, sz = [], []
    for _ in range(8):
        ig.step(1)
        times.append(ig.time())
        sz.append(ig.expect(0, SIGMA["z"]).real)
        assert abs(ig.norm2() - 1.0) < 1e-12
    assert list(ig.hranks()) == schmidt_ranks(tree, model)
    assert np.abs(np.array(sz) - exact_sigma(model, psi0, times)).max() < 1e-12


@pytest.mark.parametrize("name,steps,bounds", EXACT_GROWTH)
def test_gpu_reproduces_reference_bond_growth(hip, name, steps, bounds):
    # bond capacity 64 (the device's largest mode dimension) in place of the run's 192: the
    # oracle gives the identical run with either, since no bond grows past 17 in these steps
    print(f"[gpu expansion reference run] {name}, {steps} steps", flush=True)
    check_reference_growth(hip, name, steps, bounds, cap=64)


def test_gpu_debye_growth_tracks_reference(hip):
    print("[gpu expansion reference run] ohmic_debye, 3 steps", flush=True)
    check_debye_growth(hip, cap=64)


@pytest.mark.parametrize("n,d,r,cap", [(8, 6, 2, 6), (16, 8, 3, 12)])
def test_gpu_expansion_matches_oracle(hip, orc, n, d, r, cap):
    print(f"[gpu expansion vs oracle] modes {n} boson {d} bond {r} -> {cap}", flush=True)
    model = sb_model(n, d, r, nspf_cap=cap, nspf_lower_cap=cap)
    g, ys = side_by_side(hip, orc, model, 8)
    assert max(g.hranks()) > r + 2  # the bonds grew
    assert ys.exact_steps >= 2


def test_gpu_expansion_on_a_team_matches_oracle(hip, orc):
    print("[gpu expansion team] modes 16 boson 8 bond 4 -> 16, team of 4", flush=True)
    model = sb_model(16, 8, 4, nspf_cap=16, nspf_lower_cap=16)
    g, ys = side_by_side(hip, orc, model, 6, team_size=4)
    assert max(g.hranks()) > 6
    assert ys.exact_steps >= 2


def test_gpu_c1_expansion_grows_to_capacity_like_the_oracle(hip):
    fx = np.load(os.path.join(GOLDEN, "c1x_expansion.npz"))
    nsteps = len(fx["t"])
    print(f"[gpu expansion c1] 128 modes, boson 16, bond 8 -> 32, {nsteps} steps", flush=True)
    inst = c1_expansion_instance()
    tree, sop = sb_problem(hip, inst.model)
    ig = Integrator(hip, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=float(fx["krylov_tol"]),
                    mode=MODE_SUBSPACE_EXPANSION, **C1_EXPANSION)
    ig.set_sb_initial_state(inst.model)
    ys = RoundingYardstick()
    npert = fx["hranks_p"].shape[0]
    for k in range(nsteps):
        ig.step(1)
        assert abs(ig.time() - fx["t"][k]) < 1e-12
        ref = (fx["hranks"][k], fx["counters"][k], fx["sigma"][k])
        ps = [(fx["hranks_p"][p, k], fx["counters_p"][p, k], fx["sigma_p"][p, k]) for p in range(npert)]
        ys.check(k, _state(ig), ref, ps, fx["fragile"][k] == 0)
        # the Krylov tolerance (1e-12 per propagation) lets the oracle's norm drift by 1.5e-12 in 50 steps
        assert abs(ig.norm2() - fx["norm2"][k]) < 1e-12
    print(f"exact for {ys.exact_steps} of {nsteps} steps; largest differences (bonds, sum, Nonesite, Ntwosite, "
          f"sigma) {ys.worst}, perturbed oracle {ys.env} {ys.sigma_env:.1e}")
    grew = np.nonzero(fx["hranks"].sum(axis=1) > fx["hranks"][0].sum())[0]
    assert ys.exact_steps > grew[0]  # exact through the first step that grows bonds (step 10)
    assert max(ig.hranks()) == 32 == fx["hranks"][-1].max()
    ig.destroy()


def test_gpu_no_spawn_equals_fixed_step(hip):
    check_no_spawn_equals_fixed_step(hip)
