"""Leaves holding several bath modes: partition_modes and the harmonic mode combinations of
the reference's driver (src/mode_data.hpp:136-361, 818-886; src/spin_boson.cpp:418-441),
through the oracle library (the host code is shared with the HIP library).

Pinned by:
* partitions worked out by hand from the greedy rule for the reference inputs' keys
  (ecut 3000, high frequency bound 25, target dimension 1000, maximum dimensions 100 / 30),
  and, on random frequencies, the rule's defining properties (blocks cover the modes in
  order; a block of several modes fits the bound; adding the next mode would not);
* the block operators against Kronecker sums of textbook single-mode operators
  (H0 = sum_k w_k n_k, Hc = sum_k g_k (a_k + a_k^dagger)) and a hand-derived CSR row order;
* full-rank trees with two-mode leaves against the dense propagator exp(-iHt).
"""
import numpy as np
import pytest

from helpers import exact_sigma, make_integrator, ttns_vector
from ttns_lib_amd.capi import SIGMA, SpinBoson, TTNError, driver_blocks, mode_block, partition_modes

REF_KEYS = dict(ecut=3000.0, wc=50.0, nmax=1000, nmax4=250, nmax2=100, nmax3=30)


def _nper(w, ecut, nmax):
    return min(nmax, int(ecut / abs(w)))


def test_partition_reference_input_keys(orc):
    # low modes keep 100 or 75 states (blocks of one: 100^2 > 1000); above 2 x 25 the cap is
    # 30 states and 250 per block: w = 300, 310 keep 10 each (100 <= 250), w = 320, 330 keep
    # 9 (81, then 729 > 250), w = 490, 495, 500 keep 6 (216 <= 250, 1296 > 250), the last
    # mode is alone
    wk = [1, 2, 10, 40, 60, 100, 300, 310, 320, 330, 490, 495, 500, 505]
    assert partition_modes(orc, wk, **REF_KEYS) == [1, 1, 1, 1, 1, 1, 2, 2, 3, 1]


def test_partition_edge_cases(orc):
    # a block stops at the end of the bath: 7^2 = 49, 7^3 > 250
    assert partition_modes(orc, [400.0] * 3, **REF_KEYS) == [2, 1]
    # one Fock state per mode: every product is 1, so the rest of the bath is one block
    assert partition_modes(orc, [2000.0, 2500.0, 2900.0], **REF_KEYS) == [3]
    # a single mode above the bound still forms a block
    assert partition_modes(orc, [1.0], 3000.0, 50.0, 10, 250, 100, 30) == [1]
    # the high-frequency test is signed: two modes at +100 keep 30 states each under the high
    # caps (900 > 250, two blocks); at -100 they fall under the low caps (900 <= 1000, one block)
    assert partition_modes(orc, [100.0, 100.0], **REF_KEYS) == [1, 1]
    assert partition_modes(orc, [-100.0, -100.0], **REF_KEYS) == [2]
    assert partition_modes(orc, [-1.0, 2.0, -600.0, -700.0], **REF_KEYS) == [1, 1, 2]
    with pytest.raises(TTNError):
        partition_modes(orc, [2.0, 1.0], **REF_KEYS)
    # a bound near the int64 range: the state count saturates instead of overflowing, and
    # 100^9 = 1e18 <= 4e18 < 100^10 gives blocks of nine modes
    huge = dict(REF_KEYS, nmax=4 * 10 ** 18)
    assert partition_modes(orc, [1.0] * 20, **huge) == [9, 9, 2]
    # the largest int64 bound: 100^9 <= 2^63 - 1 < 100^10, still blocks of nine
    top = dict(REF_KEYS, nmax=2 ** 63 - 1)
    assert partition_modes(orc, [1.0] * 20, **top) == [9, 9, 2]


@pytest.mark.parametrize("seed", range(6))
def test_partition_is_the_maximal_greedy_split(orc, seed):
    rng = np.random.default_rng(seed)
    wk = np.sort(rng.uniform(20.0, 900.0, int(rng.integers(5, 60))))
    keys = dict(ecut=3000.0, wc=float(rng.choice([50.0, 200.0])), nmax=int(rng.choice([300, 1000])), nmax4=250,
                nmax2=100, nmax3=int(rng.choice([10, 30])))
    part = partition_modes(orc, wk, **keys)
    assert sum(part) == len(wk) and min(part) >= 1
    at = 0
    for b, size in enumerate(part):
        high = wk[at] > keys["wc"]
        nper = _nper(wk[at], keys["ecut"], keys["nmax3"] if high else keys["nmax2"])
        bound = keys["nmax4"] if high else keys["nmax"]
        if size > 1:
            assert nper ** size <= bound
        if at + size < len(wk):
            assert nper ** (size + 1) > bound
        at += size


def _textbook(w, g, d):
    a = np.diag(np.sqrt(np.arange(1, d)), 1)
    return np.diag(w * np.arange(d)), g * (a + a.T)


def _dense_csr(blk):
    n = len(blk["h0"])
    M = np.zeros((n, n))
    for i in range(n):
        for p in range(blk["rowptr"][i], blk["rowptr"][i + 1]):
            M[i, blk["colind"][p]] += blk["vals"][p]
    return M


@pytest.mark.parametrize("w,g,ecut,nmax", [
    ([0.7], [0.3], 3000.0, 6),
    ([0.5, 0.9], [0.2, -0.4], 3000.0, 4),
    ([0.5, 1.0, 1.6], [0.1, 0.25, 0.5], 3.3, 3),  # cutoffs 3, 3, 2
])
def test_block_operators_are_kronecker_sums(orc, w, g, ecut, nmax):
    dims = [_nper(x, ecut, nmax) for x in w]
    blk = mode_block(orc, w, g, ecut, nmax)
    n = int(np.prod(dims))
    assert blk["occupations"].shape == (n, len(w))
    # lexicographic order, mode 0 most significant
    assert np.array_equal(blk["occupations"], np.array(list(np.ndindex(*dims))).reshape(n, len(w)))
    H0 = np.zeros((n, n))
    Hc = np.zeros((n, n))
    for k, d in enumerate(dims):
        h0, hc = _textbook(w[k], g[k], d)
        left, right = np.eye(int(np.prod(dims[:k]))), np.eye(int(np.prod(dims[k + 1:])))
        H0 += np.kron(np.kron(left, h0), right)
        Hc += np.kron(np.kron(left, hc), right)
    assert np.allclose(np.diag(blk["h0"]), H0, atol=1e-15, rtol=0)
    assert np.allclose(_dense_csr(blk), Hc, atol=1e-15, rtol=0)
    # every nonzero of Hc appears exactly once
    assert len(blk["vals"]) == np.count_nonzero(Hc)


def test_block_csr_row_order(orc):
    # state (1, 1, 1) of three 3-state modes is row 13: lowering entries in mode order
    # (13 - 9, 13 - 3, 13 - 1), then raising entries in reverse mode order (13 + 1, 13 + 3, 13 + 9)
    g = [0.1, 0.2, 0.3]
    blk = mode_block(orc, [1.0, 1.0, 1.0], g, 3000.0, 3)
    lo, hi = blk["rowptr"][13], blk["rowptr"][14]
    assert list(blk["colind"][lo:hi]) == [4, 10, 12, 14, 16, 22]
    s2 = np.sqrt(2.0)
    assert np.allclose(blk["vals"][lo:hi], [g[0], g[1], g[2], s2 * g[2], s2 * g[1], s2 * g[0]], atol=1e-16)


def test_block_rejects_modes_without_states(orc):
    with pytest.raises(TTNError):
        mode_block(orc, [10.0], [0.1], 5.0, 4)  # floor(5 / 10) = 0 states
    with pytest.raises(TTNError):
        mode_block(orc, [0.0], [0.1], 5.0, 4)


def test_driver_blocks_choose_the_cap_by_first_frequency(orc):
    sizes, nmax = driver_blocks(orc, [1.0, 300.0, 310.0], ecut=3000.0, high_frequency_bound=25.0,
                                target_dimension=1000, maximum_dimension=100, maximum_dimension_high_frequency=30)
    assert sizes == [1, 2] and nmax == [100, 30]


def per_mode_model(model: SpinBoson) -> SpinBoson:
    """The same bath with one mode per leaf and each mode's block cutoff."""
    dims, at = [], 0
    for size, cap in zip(model.block_sizes, model.block_nmax):
        dims += [_nper(x, model.ecut, cap) for x in model.wk[at:at + size]]
        at += size
    return SpinBoson(wk=model.wk, gk=model.gk, boson_dim=dims, hsys=model.hsys, scoup=model.scoup, psi0=model.psi0)


BLOCK_MODELS = {
    "two pairs": dict(wk=[0.5, 0.9, 1.3, 1.8], gk=[0.3, 0.4, 0.5, 0.6], block_sizes=[2, 2], block_nmax=[3, 3],
                      ecut=3000.0, nspf=81, nspf_lower=9),
    "mixed cutoffs": dict(wk=[0.5, 1.0, 1.6], gk=[0.3, 0.45, 0.6], block_sizes=[2, 1], block_nmax=[3, 3],
                          ecut=3.3, nspf=18, nspf_lower=9),
}


@pytest.mark.parametrize("name", list(BLOCK_MODELS))
def test_full_rank_block_tree_matches_dense_propagator(orc, name):
    model = SpinBoson(hsys=np.array([[0.2, 1.0], [1.0, -0.2]], dtype=complex), **BLOCK_MODELS[name])
    tree, sop, integ = make_integrator(orc, model, dt=0.05, krylov_dim=16, krylov_tol=1e-13)
    psi0 = ttns_vector(integ, tree)
    times, sz = [], []
    for _ in range(6):
        integ.step(1)
        times.append(integ.time())
        sz.append(integ.expect(0, SIGMA["z"]).real)
    ref = exact_sigma(per_mode_model(model), psi0, times)
    assert np.abs(np.array(sz) - ref).max() < 1e-12
