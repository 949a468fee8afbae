"""GPU parity: the HIP library against the oracle (same inputs, same algorithms) and
against exact answers.  All through the C ABI."""
import numpy as np
import pytest
import scipy.linalg as sla

from helpers import exact_sigma, full_hamiltonian, make_integrator, sb_model, ttns_vector
from ttns_lib_amd.capi import SIGMA, Batch, SpinBoson, TTNError

pytestmark = pytest.mark.gpu


def _pair(hip, orc, model, **kw):
    _, _, g = make_integrator(hip, model, **kw)
    _, _, o = make_integrator(orc, model, **kw)
    return g, o


# Tolerances sit just above each case's own sensitivity: perturbing one node of the
# 8-mode tree by 1e-16 moves the oracle's <sigma> by up to 6e-12 within 12 steps,
# because Krylov truncations at tolerance 1e-12 then stop at a different dimension.
@pytest.mark.parametrize("n,d,r,steps,tol", [(1, 4, 4, 6, 1e-13), (2, 3, 3, 8, 1e-13), (4, 4, 4, 12, 1e-13),
                                            (8, 6, 4, 12, 2e-11)])
def test_small_trees_match_oracle(hip, orc, n, d, r, steps, tol):
    model = sb_model(n, d, r)
    g, o = _pair(hip, orc, model)
    for _ in range(steps):
        g.step(1)
        o.step(1)
        for op in "xyz":
            assert abs(g.expect(0, SIGMA[op]) - o.expect(0, SIGMA[op])) < tol
        assert abs(g.norm2() - o.norm2()) < tol


def test_full_rank_tree_matches_dense_propagator(hip):
    d, nm = 3, 4
    model = SpinBoson(wk=np.linspace(0.5, 2.0, nm), gk=np.linspace(0.3, 0.6, nm), boson_dim=[d] * nm, nspf=d * d,
                      nspf_lower=d, hsys=np.array([[0.2, 1.0], [1.0, -0.2]], dtype=complex))
    tree, _, g = make_integrator(hip, model, dt=0.05, krylov_dim=16, krylov_tol=1e-13)
    psi0 = ttns_vector(g, tree)
    times, sz = [], []
    for _ in range(8):
        g.step(1)
        times.append(g.time())
        sz.append(g.expect(0, SIGMA["z"]).real)
    assert np.abs(np.array(sz) - exact_sigma(model, psi0, times)).max() < 1e-12
    psi = sla.expm(-1j * full_hamiltonian(model) * times[-1]) @ psi0
    assert np.abs(ttns_vector(g, tree) - psi).max() < 1e-12


def test_c0_trajectory_matches_oracle(hip, orc):
    """Plain small config: 16 modes, boson dim 10, binary tree depth 5, bond dim 8,
    dt 0.05, 200 steps, Krylov dim 12.  Tolerance |d<sigma_z>| <= 1e-10: perturbing one
    leaf of the oracle's own initial state by 1e-15 moves <sigma_z> by up to 5e-11 over
    the same run, so no two implementations can agree more closely than that."""
    model = sb_model(16, 10, 8, alpha=0.1, wc=5.0)
    g, o = _pair(hip, orc, model, dt=0.05, krylov_dim=12)
    g.lib._call("set_observable", g.h, 0, _op(SIGMA["z"]))
    worst = 0.0
    for _ in range(4):
        g.step(50)
        o.step(50)
    traj = _trajectory(g)
    assert traj.shape == (200, 2)
    # compare the final state and, on the way, norm conservation
    assert abs(g.expect(0, SIGMA["z"]) - o.expect(0, SIGMA["z"])) < 1e-10
    assert np.abs(traj[:, 1] - 1.0).max() < 1e-12
    worst = max(worst, abs(traj[-1, 0] - o.expect(0, SIGMA["z"]).real))
    assert worst < 1e-10


def test_c0_every_step_matches_oracle(hip, orc):
    model = sb_model(16, 10, 8, alpha=0.1, wc=5.0)
    g, o = _pair(hip, orc, model, dt=0.05, krylov_dim=12)
    for _ in range(60):
        g.step(1)
        o.step(1)
        assert abs(g.expect(0, SIGMA["z"]).real - o.expect(0, SIGMA["z"]).real) < 1e-10


def test_batch_equals_individual_runs(hip):
    models = [sb_model(8, 6, 4, seed=s, alpha=0.05 * (s + 1)) for s in range(5)]
    singles, batched = [], []
    for m in models:
        _, _, a = make_integrator(hip, m)
        _, _, b = make_integrator(hip, m)
        singles.append(a)
        batched.append(b)
    for a in singles:
        a.step(6)
    B = Batch(hip, batched)
    B.step(6)
    for a, b in zip(singles, batched):
        # same kernel code per workgroup: bit-identical
        for nd in range(len(a.shapes)):
            assert np.array_equal(a.get_node(nd), b.get_node(nd))


def test_uneven_batch_matches_oracle(hip, orc):
    specs = [(4, 4, 4), (8, 6, 4), (16, 5, 6), (6, 8, 8)]
    models = [sb_model(n, d, r, seed=i) for i, (n, d, r) in enumerate(specs)]
    gs, os_ = [], []
    for m in models:
        _, _, g = make_integrator(hip, m)
        _, _, o = make_integrator(orc, m)
        gs.append(g)
        os_.append(o)
    B = Batch(hip, gs)
    for _ in range(3):
        B.step(2)
        for o in os_:
            o.step(2)
        for g, o in zip(gs, os_):
            assert abs(g.expect(0, SIGMA["z"]) - o.expect(0, SIGMA["z"])) < 1e-12


def test_nan_state_reports_numeric_error(hip):
    model = sb_model(4, 4, 4)
    _, _, g = make_integrator(hip, model)
    bad = g.get_node(4).copy()
    bad[0, 0] = np.nan
    g.set_node(4, bad)
    with pytest.raises(TTNError) as e:
        g.step(1)
    assert e.value.code == 1


def test_batch_isolates_a_nan_instance(hip):
    # one poisoned instance ends in TTN_ERR_NUMERIC naming it; every other instance of the
    # launch completes its steps bit for bit as when stepped alone
    print("[batch nan isolation] 5 instances, instance 2 poisoned", flush=True)
    models = [sb_model(8, 6, 4, seed=s, alpha=0.05 * (s + 1)) for s in range(5)]
    singles = [make_integrator(hip, m)[2] for m in models]
    batched = [make_integrator(hip, m)[2] for m in models]
    bad = batched[2].get_node(4).copy()
    bad[0, 0] = np.nan
    batched[2].set_node(4, bad)
    for k, a in enumerate(singles):
        if k != 2:
            a.step(3)
    B = Batch(hip, batched)
    with pytest.raises(TTNError) as e:
        B.step(3)
    assert e.value.code == 1 and "batch instance 2" in str(e.value)
    for k, (a, b) in enumerate(zip(singles, batched)):
        if k == 2:
            assert b.counters()[1] == 0  # the failing instance completed no step
            continue
        assert b.counters()[1] == 3 and b.time() == a.time()
        for nd in range(len(a.shapes)):
            assert np.array_equal(a.get_node(nd), b.get_node(nd))


def test_several_observables_recorded_on_device(hip, orc):
    # sigma_x, sigma_y, sigma_z, 1 on the spin and the occupation of bath mode 3, recorded after
    # every step of one launch, against the oracle's expectation values step by step
    print("[observables] 5 operators over 6 steps", flush=True)
    model = sb_model(8, 6, 4)
    _, _, g = make_integrator(hip, model)
    _, _, o = make_integrator(orc, model)
    num = np.diag(np.arange(6)).astype(complex)
    obs = [(0, SIGMA["x"]), (0, SIGMA["y"]), (0, SIGMA["z"]), (0, np.eye(2)), (4, num)]
    g.set_observables(obs)
    g.step(6)
    vals, norms = g.observables()
    assert vals.shape == (6, 5) and norms.shape == (6,)
    for k in range(6):
        o.step(1)
        want = np.array([o.expect(m, op) for m, op in obs])
        assert np.abs(vals[k] - want).max() < 1e-10, k
        assert abs(norms[k] - o.norm2()) < 1e-12
    assert np.abs(vals[-1] - np.array([g.expect(m, op) for m, op in obs])).max() < 1e-14
    # the one-observable trajectory reads the first slot
    assert np.array_equal(g.trajectory()[:, 0], vals[:, 0].real)
    # a new set clears the records
    g.set_observables(obs[2:3])
    g.step(2)
    assert g.observables()[0].shape == (2, 1)


def _op(a):
    import ctypes as C

    buf = np.ascontiguousarray(np.asarray(a, dtype=complex)).view(np.float64)
    _op.keep = buf
    return buf.ctypes.data_as(C.POINTER(C.c_double))


def _trajectory(g):
    import ctypes as C

    n = C.c_int64(0)
    g.lib._call("trajectory", g.h, C.byref(n), None)
    out = np.zeros(2 * n.value)
    g.lib._call("trajectory", g.h, C.byref(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out.reshape(-1, 2)
