"""Writes tests/golden/<name>_oracle.json: the oracle's <sigma_x,y,z>(t) and <psi|psi> over the
first steps of a single instance with the fixed-step integrator: c1 = the single larger run's
tree (Debye bath alpha 0.5, wc 2, beta 10 folded into 128 modes, boson dim 16, bond dim 32,
dt 0.025; subspace expansion is not built), c4 = the gradient configuration's tree (Ohmic alpha
0.2, wc 5, T = 0, 64 modes, boson dim 12, bond dim 24, dt 0.05).  Minutes per step on one CPU
core, so computed once here.

Usage: python tests/golden/make_single_oracle.py [nsteps] [c1|c4]   (default 2 c1)
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ttns_lib_amd.capi import SIGMA, Integrator, Library, sb_problem  # noqa: E402
from ttns_lib_amd.workloads import c1_instance, c4_instance  # noqa: E402

if __name__ == "__main__":
    nsteps = int(sys.argv[1]) if len(sys.argv) > 1 else 2
    name = sys.argv[2] if len(sys.argv) > 2 else "c1"
    OUT = os.path.join(ROOT, "tests", "golden", f"{name}_oracle.json")
    orc = Library(os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so"), "orc_")
    inst = {"c1": c1_instance, "c4": c4_instance}[name]()
    tree, sop = sb_problem(orc, inst.model)
    ig = Integrator(orc, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-12)
    ig.set_sb_initial_state(inst.model)
    rec = {"steps": [], "krylov_dim": inst.krylov_dim, "krylov_tol": 1e-12, "dt": inst.dt}
    for s in range(nsteps):
        t0 = time.time()
        ig.step(1)
        sig = {o: complex(ig.expect(0, SIGMA[o])).real for o in "xyz"}
        rec["steps"].append({"t": ig.time(), **sig, "norm2": ig.norm2(), "nh": ig.counters()[0]})
        print(f"step {s + 1}: {rec['steps'][-1]} ({time.time() - t0:.0f}s)", flush=True)
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print(f"wrote {OUT}")
