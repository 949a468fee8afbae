"""Writes tests/golden/c3_oracle.json: the oracle's <sigma_x,y,z>(t) and <psi|psi> over the
first steps of selected instances of the batched workload (full size, complex128).  The GPU
tests compare the HIP library against these numbers; the oracle needs minutes per step at
this size, so they are computed once here.

Usage: python tests/golden/make_c3_oracle.py [instance ...]   (default: 0)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ttns_lib_amd.capi import SIGMA, Integrator, Library, sb_problem  # noqa: E402
from ttns_lib_amd.workloads import c3_instances  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden", "c3_oracle.json")
NSTEPS = 2

if __name__ == "__main__":
    picks = [int(a) for a in sys.argv[1:]] or [0]
    orc = Library(os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so"), "orc_")
    ins = c3_instances()
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for k in picks:
        inst = ins[k]
        tree, sop = sb_problem(orc, inst.model)
        ig = Integrator(orc, tree, sop, dt=inst.dt, krylov_dim=inst.krylov_dim, krylov_tol=1e-12)
        ig.set_sb_initial_state(inst.model)
        rec = {"steps": [], "krylov_dim": inst.krylov_dim, "krylov_tol": 1e-12}
        for s in range(NSTEPS):
            t0 = time.time()
            ig.step(1)
            sig = {o: complex(ig.expect(0, SIGMA[o])).real for o in "xyz"}
            rec["steps"].append({"t": ig.time(), **sig, "norm2": ig.norm2(), "nh": ig.counters()[0]})
            print(f"instance {k} step {s + 1}: {rec['steps'][-1]} ({time.time() - t0:.0f}s)", flush=True)
        data[str(k)] = rec
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print(f"wrote {OUT}")
