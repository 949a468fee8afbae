"""GPU: leaves holding several bath modes (the reference driver's mode combinations),
against the dense propagator on full-rank trees and against the oracle on a partitioned
ohmic bath."""
import numpy as np
import pytest

from helpers import exact_sigma, make_integrator, ohmic_modes, ttns_vector
from test_mode_blocks import BLOCK_MODELS, per_mode_model
from ttns_lib_amd.capi import SIGMA, SpinBoson, driver_blocks, mode_block, partition_modes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(BLOCK_MODELS))
def test_full_rank_block_tree_matches_dense_propagator(hip, name):
    model = SpinBoson(hsys=np.array([[0.2, 1.0], [1.0, -0.2]], dtype=complex), **BLOCK_MODELS[name])
    tree, sop, integ = make_integrator(hip, model, dt=0.05, krylov_dim=16, krylov_tol=1e-13)
    psi0 = ttns_vector(integ, tree)
    times, sz = [], []
    for _ in range(6):
        integ.step(1)
        times.append(integ.time())
        sz.append(integ.expect(0, SIGMA["z"]).real)
    ref = exact_sigma(per_mode_model(model), psi0, times)
    assert np.abs(np.array(sz) - ref).max() < 1e-12


def _partitioned_bath(lib):
    # 16 ohmic modes: below 2 x 2.0 pairs of 10-state modes (100 states per leaf), above it
    # triples of 6-state modes (216 states per leaf)
    w, g = ohmic_modes(16, 0.1, 5.0)
    order = np.argsort(np.abs(w))
    w, g = np.asarray(w)[order], np.asarray(g)[order]
    sizes, nmax = driver_blocks(lib, w, ecut=3000.0, high_frequency_bound=2.0, target_dimension=100,
                                maximum_dimension=10, maximum_dimension_high_frequency=6)
    return SpinBoson(wk=w, gk=g, block_sizes=sizes, block_nmax=nmax, nspf=8, nspf_lower=6)


def test_partitioned_bath_matches_oracle(hip, orc):
    """The oracle's own trajectory moves by 5e-13 in 4 steps when its initial state is
    perturbed by 1e-15 relative; 1e-10 leaves room for the GPU's different summation order."""
    m_hip, m_orc = _partitioned_bath(hip), _partitioned_bath(orc)
    assert list(m_hip.block_sizes) == [2, 2, 3, 3, 3, 3] == list(m_orc.block_sizes)
    _, _, a = make_integrator(hip, m_hip)
    _, _, o = make_integrator(orc, m_orc)
    assert (100, 6) in a.shapes and (216, 6) in a.shapes
    for _ in range(4):
        a.step(1)
        o.step(1)
        for op in "xyz":
            assert abs(a.expect(0, SIGMA[op]) - o.expect(0, SIGMA[op])) < 1e-10
        assert abs(a.norm2() - o.norm2()) < 1e-12


def test_block_helpers_exported_by_the_hip_library(hip, orc):
    wk = [1, 2, 10, 40, 60, 100, 300, 310, 320, 330, 490, 495, 500, 505]
    keys = dict(ecut=3000.0, wc=50.0, nmax=1000, nmax4=250, nmax2=100, nmax3=30)
    assert partition_modes(hip, wk, **keys) == partition_modes(orc, wk, **keys)
    a, b = mode_block(hip, [0.5, 0.9], [0.2, 0.3], 3000.0, 4), mode_block(orc, [0.5, 0.9], [0.2, 0.3], 3000.0, 4)
    for k in a:
        assert np.array_equal(a[k], b[k])


def test_batch_with_block_leaves_equals_individual_runs(hip):
    # a partitioned bath and a one-mode-per-leaf bath in one launch: each instance runs alone
    # on its workgroup (team of one), so the batch gives the individual results bit for bit
    from helpers import sb_model
    from ttns_lib_amd.capi import Batch

    models = [_partitioned_bath(hip), sb_model(8, 6, 4)]
    singles = [make_integrator(hip, m, team_size=1)[2] for m in models]
    batched = [make_integrator(hip, m, team_size=1)[2] for m in models]
    for a in singles:
        a.step(3)
    B = Batch(hip, batched)
    try:
        B.step(3)
    finally:
        B.destroy()
    for a, b in zip(singles, batched):
        for nd in range(len(a.shapes)):
            assert np.array_equal(a.get_node(nd), b.get_node(nd))


def test_block_leaves_spread_over_a_team_match_oracle(hip, orc):
    # leaves of 216 states x 10 functions (2160 entries) reach the team threshold: their
    # H-actions, QRs and environment products are spread over a team of 4 workgroups
    m_hip, m_orc = _partitioned_bath(hip), _partitioned_bath(orc)
    for m in (m_hip, m_orc):
        m.nspf, m.nspf_lower = 10, 10
    _, _, a = make_integrator(hip, m_hip, team_size=4)
    _, _, o = make_integrator(orc, m_orc)
    assert (216, 10) in a.shapes
    for _ in range(3):
        a.step(1)
        o.step(1)
        for op in "xyz":
            assert abs(a.expect(0, SIGMA[op]) - o.expect(0, SIGMA[op])) < 1e-10
        assert abs(a.norm2() - o.norm2()) < 1e-12
