"""GPU: the small-vector propagator (Krylov basis in registers, operators and intermediates
in LDS, one workgroup) against the general Krylov path and against the oracle."""
import numpy as np
import pytest

from helpers import make_integrator, sb_model
from ttns_lib_amd.capi import SIGMA, Library

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def general(hip):
    """A second context on the same GPU that runs every propagation on the general path."""
    lib = Library(hip.path)
    lib.set_small(0)
    yield lib
    lib.close()


def _run_pair(hip, general, orc, model, steps, kdim=12, team=None, tol_sz=2e-11):
    kw = {} if team is None else {"team_size": team}
    _, _, s = make_integrator(hip, model, krylov_dim=kdim, **kw)
    _, _, g = make_integrator(general, model, krylov_dim=kdim, **kw)
    _, _, o = make_integrator(orc, model, krylov_dim=kdim)
    for _ in range(steps):
        s.step(1)
        g.step(1)
        o.step(1)
        for op in "xyz":
            es, eg, eo = (ig.expect(0, SIGMA[op]) for ig in (s, g, o))
            assert abs(es - eg) < tol_sz
            assert abs(es - eo) < tol_sz
        assert abs(s.norm2() - o.norm2()) < 1e-12
    return s, g


@pytest.mark.parametrize("kdim", [12, 16])
def test_small_tree_every_propagation_small(hip, general, orc, kdim):
    # every node tensor and bond matrix of this tree has at most 512 entries
    s, g = _run_pair(hip, general, orc, sb_model(8, 6, 4), 8, kdim=kdim)
    # same algorithm and step control: the H-application counts agree up to sub-steps whose
    # error estimate lands on the tolerance within rounding
    ns, ng = s.counters()[0], g.counters()[0]
    assert abs(ns - ng) <= 0.01 * ng
    assert s.profile()["n_small"] > 0
    assert g.profile()["n_small"] == 0


def test_19_level_leaves_small(hip, general, orc):
    # 19-level bath modes, strong coupling: leaf tensors of 19 x 6 on the small propagator
    # (2433 H applications in 4 steps: rounding differences of the GPU's contractions reach
    # a few 1e-10 on both GPU paths)
    s, _ = _run_pair(hip, general, orc, sb_model(8, 19, 6, wc=5.0, alpha=0.5), 4, tol_sz=1e-9)
    assert s.profile()["n_small"] > 0


def _perturbed(integ, seed, rel=1e-15):
    rng = np.random.default_rng(seed)
    for nd in range(len(integ.shapes)):
        a = integ.get_node(nd)
        integ.set_node(nd, a * (1 + rel * rng.standard_normal(a.shape)))


def test_stiff_leaves_small(hip, general, orc):
    """Frequencies up to ~4 wc = 80 with mostly unoccupied initial bond spaces: the one-site
    step fills the unoccupied directions from rounding noise, so trajectories of any two
    implementations differ far above rounding.  The bound is measured here: the spread of the
    oracle's own trajectory when its initial state is perturbed by 1e-15 relative."""
    m = sb_model(8, 19, 6, wc=20.0, alpha=0.5)
    steps = 2
    runs = {}
    for name, lib, pert in (("small", hip, 0), ("general", general, 0), ("orc", orc, 0), ("orc1", orc, 1),
                            ("orc2", orc, 2)):
        ig = make_integrator(lib, m)[2]
        if pert:
            _perturbed(ig, pert)
        traj = []
        for _ in range(steps):
            ig.step(1)
            traj.append([ig.expect(0, SIGMA[op]) for op in "xyz"] + [ig.norm2()])
        runs[name] = (ig, np.array(traj))
    o = runs["orc"][1]
    spread = max(np.abs(runs[k][1][:, :3] - o[:, :3]).max() for k in ("orc1", "orc2"))
    assert spread < 1e-4
    for k in ("small", "general"):
        t = runs[k][1]
        assert np.abs(t[:, :3] - o[:, :3]).max() < 10 * spread + 1e-10, k
        assert np.abs(t[:, 3] - o[:, 3]).max() < 1e-12, k
    assert runs["small"][0].profile()["n_small"] > 0


def test_small_leaves_under_team(hip, general, orc):
    # 16-mode bond-16 tree on a team of 8: leaves (and their bonds) on the small propagator,
    # the upper nodes spread over the team
    s, _ = _run_pair(hip, general, orc, sb_model(16, 10, 16), 3, team=8, tol_sz=1e-10)
    assert s.profile()["n_small"] > 0


def test_small_max_rejects_negative(hip):
    lib = Library(hip.path)
    try:
        with pytest.raises(Exception):
            lib.set_small(-1)
        lib.set_small(100)
    finally:
        lib.close()


def test_node_stats_account_for_every_node_application(hip):
    # per-node H-application counts of the node propagations add up to the integrator's
    # counter, on the small propagator and on the general path alike; every Krylov size is
    # between 1 and the Krylov dimension
    for m in (sb_model(8, 6, 4), sb_model(8, 8, 24)):
        ig = make_integrator(hip, m)[2]
        ig.step(3)
        st = ig.node_stats()
        assert st[:, 0].sum() == ig.counters()[0]
        for napp, sumj in ((st[:, 0], st[:, 1]), (st[:, 2], st[:, 3])):
            used = napp > 0
            assert np.all(sumj[used] >= napp[used]) and np.all(sumj[used] <= 12 * napp[used])
        assert st[:, 2].sum() > 0
        # device ticks of the node and bond propagations where there were any
        assert np.all(st[st[:, 0] > 0, 4] > 0) and np.all(st[st[:, 2] > 0, 5] > 0)
