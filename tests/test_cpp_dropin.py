"""The C++ drop-in header (include/ttns_amd.hpp) used the way the reference's driver uses its
API (tests/cpp/sb_driver.cpp: ntree_builder topology, sop_operator::bind, httensor initial
state, the expansion integrator's step loop with the calls of src/spin_boson.cpp:41-107 --
mel(), t(), engine().Nonesite() / Ntwosite(), nh_applications(), hrank() -- and
partition_modes / harmonic_mode_combination for the leaves), plus what a caller may change
between steps: dt() (read on every call), a coefficient of the operator, reset_nsteps() and
clear() + initialise().  On the CPU the driver
runs over the oracle (tests/cpp/orc_as_ttn.cpp) and must reproduce, bit for bit, the
trajectory of the library's own spin-boson builder; on the GPU it runs over the HIP library
and must agree with the oracle within the trajectory tolerance."""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_integrator, ohmic_modes, sb_model
from ttns_lib_amd.capi import MODE_SUBSPACE_EXPANSION, SIGMA, Integrator, SpinBoson, driver_blocks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# nmodes boson_dim nspf nsteps dt alpha wc seed krylov_dim [ecut high_frequency_bound
# target_dimension maximum_dimension_high_frequency] [--option value ...]; the third case puts
# pairs of 10-state modes and triples of 6-state modes on the leaves; the others change the
# integrator's state between steps the way a caller of the reference's API can
CASES = [(8, 6, 4, 12, 0.05, 0.1, 5.0, 0, 12), (16, 10, 8, 8, 0.05, 0.3, 5.0, 3, 12),
         (16, 10, 8, 6, 0.05, 0.1, 5.0, 0, 12, 3000.0, 2.0, 100, 6),
         # bonds grow from 2 to 6 under the expansion integrator, dt halves after 3 steps
         (8, 6, 2, 6, 0.05, 0.1, 5.0, 0, 12, "--capacity", 6, "--spawn", 1e-8, "--min-unoccupied", 2,
          "--dt2", 0.025, "--dt2-at", 3),
         # the system Hamiltonian's coefficient doubles after 2 steps, nsteps restarts after 4
         (8, 6, 4, 7, 0.05, 0.1, 5.0, 1, 12, "--coeff0", 2.0, "--coeff0-at", 2, "--reset-at", 4),
         # clear() + initialise() after 3 steps (the state carries over, the step count restarts)
         (8, 6, 4, 6, 0.05, 0.1, 5.0, 2, 12, "--clear-at", 3)]
# checkpoint after 2 steps of a growing tree, back to it after 5: rows 6.. repeat rows 3..
CHECKPOINT = (8, 6, 2, 8, 0.05, 0.1, 5.0, 0, 12, "--capacity", 6, "--spawn", 1e-8, "--min-unoccupied", 2,
              "--checkpoint-at", 2, "--restore-at", 5)


def _split(case):
    pos = [x for x in case if not isinstance(x, str)]
    k = next((i for i, x in enumerate(case) if isinstance(x, str)), len(case))
    opt = {case[i][2:]: case[i + 1] for i in range(k, len(case), 2)}
    return tuple(case[:k]), opt


def _run(binary, case):
    args = [binary] + [str(x) for x in case]
    out = subprocess.run(args, check=True, capture_output=True, text=True, timeout=600).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    # t, <sigma_z>, then (after every step) Nonesite, Ntwosite, nh/size, then the hranks
    head = [[float(r[0]), float(r[1])] for r in rows]
    counts = [[int(r[2]), int(r[3]), float(r[4])] for r in rows[1:]]
    hr = [[int(x) for x in r[2:]] for r in rows[:1]] + [[int(x) for x in r[5:]] for r in rows[1:]]
    return np.array(head), np.array(counts), hr


def _oracle(orc, case):
    """The same calls on the oracle through the Python binding."""
    pos, opt = _split(case)
    n, d, r, steps, dt, alpha, wc, seed, kdim = pos[:9]
    cap = int(opt.get("capacity", r))
    if len(pos) > 9:
        ecut, hfb, target, nhigh = pos[9:]
        w, g = ohmic_modes(n, alpha, wc)
        sizes, nmax = driver_blocks(orc, w, ecut=ecut, high_frequency_bound=hfb, target_dimension=target,
                                    maximum_dimension=d, maximum_dimension_high_frequency=nhigh)
        assert max(sizes) > 1
        model = SpinBoson(wk=w, gk=g, block_sizes=sizes, block_nmax=nmax, nspf=r, nspf_lower=r, seed=seed)
    else:
        model = sb_model(n, d, r, alpha=alpha, wc=wc, seed=seed, nspf_cap=cap, nspf_lower_cap=cap)
    knobs = dict(mode=MODE_SUBSPACE_EXPANSION, spawning_threshold=opt.get("spawn", 1e-7),
                 unoccupied_threshold=opt.get("unoccupied", 1e-4), minimum_unoccupied=int(opt.get("min-unoccupied", 0)))
    tree, sop, o = make_integrator(orc, model, dt=dt, krylov_dim=kdim, **knobs)
    nodes = len(o.shapes)
    head = [[0.0, o.expect(0, SIGMA["z"]).real]]
    counts, hr = [], [list(o.hranks()[1:])]
    coeff = np.ones(1 + 2 * (tree.size()[1] - 1), dtype=complex)  # H_sys, then H0_i and S_z Hc_i per leaf
    for k in range(1, steps + 1):
        o.step(1)
        head.append([o.time(), o.expect(0, SIGMA["z"]).real])
        counts.append([*o.expansion_counters(), o.counters()[0] / nodes])
        hr.append(list(o.hranks()[1:]))
        if k == opt.get("dt2-at"):
            o.set_dt(opt["dt2"])
        if k == opt.get("coeff0-at"):
            coeff[0] *= opt["coeff0"]
            o.set_coefficients(coeff)
        if k == opt.get("reset-at"):
            o.reset_nsteps()
        if k == opt.get("clear-at"):
            state = [o.get_node(i) for i in range(nodes)]
            dtnow = o.get_dt()
            o.destroy()
            o = Integrator(orc, tree, sop, dt=dtnow, krylov_dim=kdim, krylov_tol=1e-12, **knobs)
            for i, a in enumerate(state):
                o.set_node(i, a)
    return np.array(head), np.array(counts), hr


def _driver(name):
    binary = os.path.join(ROOT, "oracle" if name == "sb_driver_orc" else "ttns_lib_amd", "_build", name)
    if name == "sb_driver_orc" and not os.path.exists(binary):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_build/sb_driver_orc"], check=True)
    assert os.path.exists(binary), "build the library first (make -C ttns_lib_amd/csrc)"
    return binary


@pytest.mark.parametrize("case", CASES)
def test_cpp_driver_over_oracle_reproduces_builtin_problem(orc, case):
    head, counts, hr = _run(_driver("sb_driver_orc"), case)
    ho, co, hro = _oracle(orc, case)
    assert np.array_equal(head, ho)
    assert np.array_equal(counts, co)
    assert hr == hro


@pytest.mark.gpu
def test_cpp_driver_checkpoint_restore_on_gpu():
    print(f"[cpp driver] checkpoint case {CHECKPOINT}", flush=True)
    head, counts, hr = _run(_driver("sb_driver"), CHECKPOINT)
    # rows after the restore (6..8) repeat the rows after the checkpoint (3..5) bit for bit
    assert np.array_equal(head[6:9], head[3:6]) and np.array_equal(counts[5:8], counts[2:5])
    assert hr[6:9] == hr[3:6] and hr[3] != hr[0]


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_cpp_driver_on_gpu_matches_oracle(orc, case):
    print(f"[cpp driver] case {case}", flush=True)
    head, counts, hr = _run(_driver("sb_driver"), case)
    ho, co, hro = _oracle(orc, case)
    assert np.abs(head[:, 0] - ho[:, 0]).max() < 1e-14
    assert np.abs(head[:, 1] - ho[:, 1]).max() < 1e-10
    assert hr == hro
    assert np.array_equal(counts[:, :2], co[:, :2])
