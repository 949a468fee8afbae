"""GPU: teams of workgroups advancing one instance (barriers in the XCD's L2, coherent
loads, reductions summed in rank order).  Against the oracle, and team against team."""
import numpy as np
import pytest

from helpers import make_integrator, sb_model
from ttns_lib_amd.capi import SIGMA, Batch, Library

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spread(hip):
    """A second context on the same GPU whose operations are all spread over the team
    (team_min = 1) and none runs on the small-vector propagator, so every team primitive
    runs even on small trees."""
    lib = Library(hip.path)
    lib.set_tuning(team_min=1, teams=True)
    lib.set_small(0)
    yield lib
    lib.close()


@pytest.mark.parametrize("team", [2, 4, 8])
def test_spread_team_matches_oracle(spread, orc, team):
    m = sb_model(8, 6, 4)
    _, _, g = make_integrator(spread, m, team_size=team)
    _, _, o = make_integrator(orc, m)
    for _ in range(8):
        g.step(1)
        o.step(1)
        # same bound as the single-workgroup path on this tree (its own sensitivity)
        for op in "xyz":
            assert abs(g.expect(0, SIGMA[op]) - o.expect(0, SIGMA[op])) < 2e-11
        assert abs(g.norm2() - o.norm2()) < 1e-12


def test_team_on_medium_tree_matches_oracle(hip, orc):
    # upper node tensors of 256 x 16 entries: spread with the default threshold, the
    # leaves and their parents run on the leader
    m = sb_model(16, 10, 16)
    _, _, g = make_integrator(hip, m, team_size=8)
    _, _, o = make_integrator(orc, m)
    for _ in range(4):
        g.step(1)
        o.step(1)
        assert abs(g.expect(0, SIGMA["z"]) - o.expect(0, SIGMA["z"])) < 1e-10
        assert abs(g.norm2() - o.norm2()) < 1e-12


def test_team_batch_equals_individual_runs(spread):
    # teams of the same size compute the same sums in the same order: bit-identical
    models = [sb_model(8, 6, 4, seed=s, alpha=0.05 * (s + 1)) for s in range(3)]
    sizes = [1, 3, 5]
    singles, batched = [], []
    for m, t in zip(models, sizes):
        singles.append(make_integrator(spread, m, team_size=t)[2])
        batched.append(make_integrator(spread, m, team_size=t)[2])
    for a in singles:
        a.step(4)
    B = Batch(spread, batched)
    B.step(4)
    assert list(B.teams()) == sizes
    for a, b in zip(singles, batched):
        for nd in range(len(a.shapes)):
            assert np.array_equal(a.get_node(nd), b.get_node(nd))


def test_autotuner_gives_large_instances_teams(hip):
    small = sb_model(4, 4, 4)
    large = sb_model(16, 10, 16)
    igs = [make_integrator(hip, small)[2], make_integrator(hip, large)[2]]
    B = Batch(hip, igs)
    B.step(1)
    t = B.teams()
    assert t[0] == 1 and t[1] > 1
    assert t.sum() <= 256


@pytest.mark.parametrize("team", [2, 8])
def test_spread_team_conserves_norm_on_medium_tree(spread, orc, team):
    """Every operation of a 16-mode, bond-16 tree spread over the team (bond matrices and
    leaves included): the propagations stay unitary to rounding and track the oracle."""
    m = sb_model(16, 10, 16)
    _, _, g = make_integrator(spread, m, team_size=team)
    _, _, o = make_integrator(orc, m)
    for _ in range(3):
        g.step(1)
        o.step(1)
        assert abs(g.norm2() - 1.0) < 1e-12
        assert abs(g.expect(0, SIGMA["z"]) - o.expect(0, SIGMA["z"])) < 1e-10
