"""GPU: the batched sweep replayed from a captured hipGraph gives, replay after replay, the
same bits as direct launches of identical instances (eager work and synchronisations in
between), with teams and without."""
import numpy as np
import pytest

from helpers import make_integrator, sb_model
from ttns_lib_amd.capi import SIGMA, Batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("teams", [[1, 1, 1], [1, 3, 5]])
def test_graph_replays_match_direct_launches(hip, teams):
    models = [sb_model(8, 6, 4, seed=s, alpha=0.05 * (s + 1)) for s in range(len(teams))]
    eager = [make_integrator(hip, m, team_size=t)[2] for m, t in zip(models, teams)]
    graph = [make_integrator(hip, m, team_size=t)[2] for m, t in zip(models, teams)]
    BE, BG = Batch(hip, eager), Batch(hip, graph)
    BE.set_graph(False)
    BG.set_graph(True)
    for call in range(5):
        BG.step(1)
        BE.step(1)
        for a in eager:  # eager work of the program between replays
            a.expect(0, SIGMA["z"])
        for a, b in zip(eager, graph):
            for nd in range(len(a.shapes)):
                assert np.array_equal(a.get_node(nd), b.get_node(nd)), (call, nd)
    captures, replays = BG.graph_stats()
    assert replays == 5
    # the first call (prepare + ladder) and the regular calls need one capture each
    assert captures == 2
    assert BE.graph_stats() == (0, 0)
