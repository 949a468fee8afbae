"""Writes tests/golden/c1x_expansion.npz: the oracle's trajectory of the single larger run with
subspace expansion on (Debye bath alpha 0.5, wc 2, beta 10 folded into 128 modes, boson dim 16,
initial bond dimension 8 grown towards capacity 32, dt 0.025, Krylov dim 12), and the same run
started from node tensors multiplied by (1 + 1e-15 N(0,1)) for three seeds.

Which functions the expansion adds is decided by comparing two-site eigenvalues and singular
values with fixed thresholds, and directions that carry no weight yet are hypersensitive to
rounding: the perturbed runs show how far two correct integrators that differ only by rounding
drift apart (bond dimensions, counters and <sigma>), which is the yardstick for the HIP path.

Arrays, per step k = 1..nsteps (p = perturbation seed 7, 8, 9):
  hranks (nsteps, nnodes), counters (nsteps, 2) [Nonesite, Ntwosite], sigma (nsteps, 3) [x, y, z],
  t (nsteps,), norm2 (nsteps,), hranks_p (3, nsteps, nnodes), counters_p (3, nsteps, 2),
  sigma_p (3, nsteps, 3), margin (nsteps,): tests/helpers.py's decision_margin of the
  unperturbed run's expansion log, fragile (nsteps,): its fragile_decisions against the
  perturbed runs' logs (a step's decisions are robust to rounding when there are none).

About 0.3-0.9 s per step per run on one CPU core.
Usage: python tests/golden/make_c1x_expansion.py [nsteps]   (default 50)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from ttns_lib_amd.capi import MODE_SUBSPACE_EXPANSION, SIGMA, Integrator, Library, sb_problem  # noqa: E402
from ttns_lib_amd.workloads import C1_EXPANSION, GOLDEN, c1_expansion_instance  # noqa: E402
from helpers import LOG_CAPACITY, decision_margin, fragile_decisions, perturb_nodes  # noqa: E402

SEEDS = (7, 8, 9)
KRYLOV_TOL = 1e-12


def make(lib, seed=None):
    inst = c1_expansion_instance()
    tree, sop = sb_problem(lib, inst.model)
    ig = Integrator(lib, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=KRYLOV_TOL,
                    mode=MODE_SUBSPACE_EXPANSION, **C1_EXPANSION)
    ig.set_sb_initial_state(inst.model)
    return perturb_nodes(ig, seed) if seed is not None else ig


def run(ig, nsteps, tag):
    hr, cnt, sig, t, nrm, logs = [], [], [], [], [], []
    for k in range(nsteps):
        t0 = time.time()
        ig.set_expansion_log(LOG_CAPACITY)
        ig.step(1)
        logs.append(ig.expansion_log())
        hr.append(list(ig.hranks()))
        cnt.append(list(ig.expansion_counters()))
        sig.append([ig.expect(0, SIGMA[q]).real for q in "xyz"])
        t.append(ig.time())
        nrm.append(ig.norm2())
        print(f"{tag} step {k + 1}: max bond {max(hr[-1])} sum {sum(hr[-1])} counters {cnt[-1]} "
              f"({time.time() - t0:.1f}s)", flush=True)
    return (np.array(hr, dtype=np.int16), np.array(cnt, dtype=np.int64), np.array(sig), np.array(t),
            np.array(nrm), logs)


if __name__ == "__main__":
    nsteps = int(sys.argv[1]) if len(sys.argv) > 1 else 50
    orc = Library(os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so"), "orc_")
    hr, cnt, sig, t, nrm, logs = run(make(orc), nsteps, "oracle")
    pert = [run(make(orc, s), nsteps, f"seed {s}") for s in SEEDS]
    spawn = C1_EXPANSION["spawning_threshold"]
    margin = np.array([decision_margin(lg, spawn) for lg in logs])
    fragile = np.array([fragile_decisions(lg, [p[5][k] for p in pert], spawn) for k, lg in enumerate(logs)])
    for k in range(nsteps):
        print(f"step {k + 1}: margin {margin[k]:.3f} fragile {fragile[k]}")
    out = os.path.join(GOLDEN, "c1x_expansion.npz")
    np.savez_compressed(out, hranks=hr, counters=cnt, sigma=sig, t=t, norm2=nrm, krylov_tol=KRYLOV_TOL,
                        margin=margin, fragile=fragile,
                        hranks_p=np.stack([p[0] for p in pert]), counters_p=np.stack([p[1] for p in pert]),
                        sigma_p=np.stack([p[2] for p in pert]), seeds=np.array(SEEDS))
    print(f"wrote {out}")
