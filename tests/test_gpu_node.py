"""GPU: the node-level entry points (ttn_node_*: full-GPU batched FP64-MFMA GEMMs) against the
oracle's H-action, mean fields and Krylov propagator on the same node (node_fixture)."""
import numpy as np
import pytest

from node_fixture import NodeProblem
from ttns_lib_amd.capi import GEMM_TILES, Node

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("d0,d1,r,nt", [(16, 16, 16, 5), (20, 12, 24, 3), (64, 64, 64, 40)])
def test_node_heff_matches_oracle(hip, orc, d0, d1, r, nt):
    print(f"[node heff] {d0}x{d1}x{r} terms {nt}", flush=True)
    P = NodeProblem(orc, d0, d1, r, nt, seed=d0 + nt)
    g = Node(hip, P.op0, P.op1, P.mf)
    g.set(P.v)
    assert _rel(g.heff(), P.heff(P.v)) < 1e-13
    # the resident input is unchanged and a second application gives the same bits
    assert np.array_equal(g.get(), P.v)
    a = g.heff()
    assert np.array_equal(a, g.heff(reps=2))
    g.destroy()


@pytest.mark.parametrize("d0,d1,r,nt", [(16, 16, 16, 5), (20, 12, 24, 3), (40, 50, 36, 7), (64, 64, 64, 40),
                                         (64, 8, 64, 3)])
def test_node_heff_forms_match_oracle(hip, orc, d0, d1, r, nt):
    """The staged H-action (three batched GEMMs) and the fused one (one kernel, per-term
    intermediates in LDS, term groups summed in order) against the oracle, each repeating itself
    bit for bit; ragged shapes exercise the zero padding of the 16 x 16 MFMA tiles."""
    print(f"[node heff forms] {d0}x{d1}x{r} terms {nt}", flush=True)
    P = NodeProblem(orc, d0, d1, r, nt, seed=d1 + nt)
    want = P.heff(P.v)
    outs = {}
    try:
        for form in (1, 2):
            hip.set_heff_form(form)
            g = Node(hip, P.op0, P.op1, P.mf)
            g.set(P.v)
            outs[form] = g.heff()
            assert np.array_equal(outs[form], g.heff()), form
            g.destroy()
    finally:
        hip.set_heff_form(0)
    for form, out in outs.items():
        assert _rel(out, want) < 1e-13, form
    assert _rel(outs[2], outs[1]) < 1e-14


@pytest.mark.parametrize("d0,d1,r,nt", [(16, 16, 16, 5), (20, 12, 24, 3), (40, 50, 36, 7), (64, 64, 64, 40),
                                         (64, 8, 64, 3)])
@pytest.mark.parametrize("child", [0, 1])
def test_node_mean_field_forms_match_oracle(hip, orc, d0, d1, r, nt, child):
    """The staged mean-field update (three batched GEMMs) and the fused one (one kernel over the
    other child's slabs, the slab's Z in LDS, chunks summed in order) against the oracle, each
    repeating itself bit for bit."""
    print(f"[node mf forms] {d0}x{d1}x{r} terms {nt} child {child}", flush=True)
    P = NodeProblem(orc, d0, d1, r, nt, seed=d0 + 2 * nt + child)
    want = P.mf_child(child)
    outs = {}
    try:
        for form in (1, 2):
            hip.set_heff_form(form)
            g = Node(hip, P.op0, P.op1, P.mf)
            g.set(P.v)
            outs[form] = g.mf(child)
            assert np.array_equal(outs[form], g.mf(child)), form
            g.destroy()
    finally:
        hip.set_heff_form(0)
    for form, out in outs.items():
        assert _rel(out, want) < 1e-13, form
    assert _rel(outs[2], outs[1]) < 1e-14


@pytest.mark.parametrize("d0,d1,r,nt", [(16, 16, 16, 5), (20, 12, 24, 3), (64, 64, 64, 40)])
@pytest.mark.parametrize("child", [0, 1])
def test_node_mean_fields_match_oracle(hip, orc, d0, d1, r, nt, child):
    print(f"[node mf] {d0}x{d1}x{r} terms {nt} child {child}", flush=True)
    P = NodeProblem(orc, d0, d1, r, nt, seed=7 + d1)
    g = Node(hip, P.op0, P.op1, P.mf)
    g.set(P.v)
    assert _rel(g.mf(child), P.mf_child(child)) < 1e-13
    g.destroy()


# the last case is the contraction microbenchmark's node (64 x 64 x 64, 40 terms) with tau chosen
# so that the propagation fills the whole 16-vector Lanczos space in one sub-step (16 H-actions)
@pytest.mark.parametrize("d0,d1,r,nt,tau,scale,na_want", [(16, 16, 16, 5, 0.05, None, None),
                                                          (24, 20, 16, 4, 0.3, None, None),
                                                          (64, 64, 64, 40, 0.12, 0.5, 16)])
def test_node_expv_matches_oracle(hip, orc, d0, d1, r, nt, tau, scale, na_want):
    print(f"[node expv] {d0}x{d1}x{r} terms {nt} tau {tau}", flush=True)
    P = NodeProblem(orc, d0, d1, r, nt, seed=11, scale=1.0 / d0 if scale is None else scale)
    v = P.v / np.linalg.norm(P.v)
    g = Node(hip, P.op0, P.op1, P.mf)
    g.set(v)
    na = g.expv(tau, krylov_dim=16, tol=1e-12)
    want, na_orc = P.expv(v, tau, 16, 1e-12)
    assert na == na_orc
    assert na_want is None or na == na_want
    assert _rel(g.get(), want) < 1e-11
    # unitary propagation
    assert abs(np.linalg.norm(g.get()) - 1.0) < 1e-11
    g.destroy()


def test_autotuner_keeps_results_and_reports_its_tiles(hip, orc):
    """The contraction microbenchmark's node with the tile autotuner off and on: the tuned
    H-action equals the oracle's, repeats itself bit for bit from the table, and the table holds
    a valid tile and K split for every GEMM shape of the H-action."""
    P = NodeProblem(orc, 64, 64, 64, 40, seed=3)
    hip.set_autotune(True)
    g = Node(hip, P.op0, P.op1, P.mf)
    g.set(P.v)
    out = g.heff()
    table = hip.autotune_table()
    print("[autotune] " + "; ".join(f"{e[:4]} tile {GEMM_TILES[e[5]]} split {e[6]} {e[7]:.3f} ms" for e in table),
          flush=True)
    assert len(table) >= 2
    assert all(0 <= e[5] < len(GEMM_TILES) and e[6] >= 1 and e[7] > 0 for e in table)
    assert _rel(out, P.heff(P.v)) < 1e-13
    assert np.array_equal(out, g.heff())
    g.destroy()
    hip.set_autotune(False)
    try:
        g0 = Node(hip, P.op0, P.op1, P.mf)
        g0.set(P.v)
        assert _rel(g0.heff(), out) < 1e-14
        g0.destroy()
    finally:
        hip.set_autotune(True)
