"""The C++ drop-in header (include/ttns_amd.hpp) used the way the reference's driver uses its
API (tests/cpp/sb_driver.cpp: ntree_builder topology, sop_operator::bind, httensor initial
state, the expansion integrator's step loop, mel(), and partition_modes /
harmonic_mode_combination for the leaves).  On the CPU the driver
runs over the oracle (tests/cpp/orc_as_ttn.cpp) and must reproduce, bit for bit, the
trajectory of the library's own spin-boson builder; on the GPU it runs over the HIP library
and must agree with the oracle within the trajectory tolerance."""
import os
import subprocess

import numpy as np
import pytest

from helpers import make_integrator, ohmic_modes, sb_model
from ttns_lib_amd.capi import SIGMA, SpinBoson, driver_blocks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# nmodes boson_dim nspf nsteps dt alpha wc seed krylov_dim [ecut high_frequency_bound
# target_dimension maximum_dimension_high_frequency]; the last case puts pairs of 10-state
# modes and triples of 6-state modes on the leaves
CASES = [(8, 6, 4, 12, 0.05, 0.1, 5.0, 0, 12), (16, 10, 8, 8, 0.05, 0.3, 5.0, 3, 12),
         (16, 10, 8, 6, 0.05, 0.1, 5.0, 0, 12, 3000.0, 2.0, 100, 6)]


def _run(binary, case):
    args = [binary] + [str(x) for x in case]
    out = subprocess.run(args, check=True, capture_output=True, text=True, timeout=600).stdout
    rows = [line.split() for line in out.strip().splitlines()]
    return np.array([float(r[0]) for r in rows]), np.array([float(r[1]) for r in rows])


def _oracle(orc, case):
    n, d, r, steps, dt, alpha, wc, seed, kdim = case[:9]
    if len(case) > 9:
        ecut, hfb, target, nhigh = case[9:]
        w, g = ohmic_modes(n, alpha, wc)
        sizes, nmax = driver_blocks(orc, w, ecut=ecut, high_frequency_bound=hfb, target_dimension=target,
                                    maximum_dimension=d, maximum_dimension_high_frequency=nhigh)
        assert max(sizes) > 1
        model = SpinBoson(wk=w, gk=g, block_sizes=sizes, block_nmax=nmax, nspf=r, nspf_lower=r, seed=seed)
    else:
        model = sb_model(n, d, r, alpha=alpha, wc=wc, seed=seed)
    _, _, o = make_integrator(orc, model, dt=dt, krylov_dim=kdim)
    t, z = [0.0], [o.expect(0, SIGMA["z"]).real]
    for _ in range(steps):
        o.step(1)
        t.append(o.time())
        z.append(o.expect(0, SIGMA["z"]).real)
    return np.array(t), np.array(z)


@pytest.mark.parametrize("case", CASES)
def test_cpp_driver_over_oracle_reproduces_builtin_problem(orc, case):
    binary = os.path.join(ROOT, "oracle", "_build", "sb_driver_orc")
    if not os.path.exists(binary):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "_build/sb_driver_orc"], check=True)
    t, z = _run(binary, case)
    to, zo = _oracle(orc, case)
    assert np.array_equal(t, to)
    assert np.array_equal(z, zo)


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_cpp_driver_on_gpu_matches_oracle(orc, case):
    binary = os.path.join(ROOT, "ttns_lib_amd", "_build", "sb_driver")
    assert os.path.exists(binary), "build the library first (make -C ttns_lib_amd/csrc)"
    print(f"[cpp driver] case {case}", flush=True)
    t, z = _run(binary, case)
    to, zo = _oracle(orc, case)
    assert np.abs(t - to).max() < 1e-14
    assert np.abs(z - zo).max() < 1e-10
