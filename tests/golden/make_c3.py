"""Writes tests/golden/c3_instances.json: the frozen 64-instance batched workload
(bath parameters, discretised (w_k, g_k), boson and bond dimensions)."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from ttns_lib_amd.workloads import C3_FILE, make_c3_spec  # noqa: E402

if __name__ == "__main__":
    spec = make_c3_spec()
    with open(C3_FILE, "w") as f:
        json.dump(spec, f, indent=0)
    print(f"wrote {len(spec)} instances to {C3_FILE}")
