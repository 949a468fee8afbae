"""Spin-boson driver over the HIP library, a drop-in for the reference's `mlmctdh.x input.json`
(src/spin_boson.cpp:182-700):

* reads the reference's JSON input.  Keys are matched with spaces removed, as its loader does
  (e.g. "krylov tolerance" -> krylovtolerance, :253).
* takes the discretised bath as w_k g_k pairs (`--bath`): a two-column file, or the log that the
  reference writes to stderr, whose pairs sit between the "w_k \\t g_k" header and the line
  "bath discretised." (:356-363).  The orthogonal-polynomial discretisation itself
  (:292-354, un-vendored `orthopol`) is not part of this library.
* builds the leaves with partition_modes and the harmonic mode blocks (:418-441, 515-538), the
  tree, the Hamiltonian and the seeded initial state.
* steps while t < tmax and writes the reference's output rows (`run`, :41-107):
  `t <O_1> ... <O_n> Nonesite Ntwosite nh/|A| [hranks of the non-root nodes]`, 16 significant
  digits, with the t = 0 row first.

Bond dimensions stay at their initial values: the bond growth of the expansion integrator is
not available on the device, so an input whose capacities exceed the initial bond dimensions
runs the one-site step (as the reference does while no bond grows) and says so on stderr.  The
expansion counters (Nonesite, Ntwosite) are written as 0.

usage: python -m ttns_lib_amd.driver input.json --bath bath.txt [--output out] [--steps N]
"""
from __future__ import annotations

import argparse
import json
import re
import sys
from typing import Optional, TextIO

import numpy as np

from .capi import Integrator, Library, SpinBoson, driver_blocks, sb_problem


def normalise_keys(doc: dict) -> dict:
    return {k.replace(" ", "").lower(): v for k, v in doc.items()}


def load_input(path: str) -> dict:
    with open(path) as f:
        text = f.read()
    # the reference's own input files end with a trailing comma before the closing brace
    return normalise_keys(json.loads(re.sub(r",\s*}\s*$", "}", text)))


def load_bath(path: str):
    """(wk, gk) from a two-column file or from the reference's stderr log."""
    with open(path, "rb") as f:
        lines = f.read().decode().split("\n")
    head = [i for i, ln in enumerate(lines) if ln.rstrip().endswith("g_k")]
    if head:
        end = next((i for i, ln in enumerate(lines) if ln.strip() == "bath discretised."), len(lines))
        lines = lines[head[0] + 1:end]
    pairs = [ln.split() for ln in lines if ln.strip() and not ln.lstrip().startswith("#")]
    wk = np.array([float(p[0]) for p in pairs])
    gk = np.array([float(p[1]) for p in pairs])
    if len(wk) == 0:
        raise ValueError(f"{path}: no w_k g_k pairs")
    return wk, gk


def _cplx(x):
    return complex(x.replace(" ", "")) if isinstance(x, str) else complex(x)


def observables(cfg: dict):
    ops = cfg["ops"]
    mats = [ops] if np.ndim(np.array(ops, dtype=object)) == 2 else ops
    return [np.array([[_cplx(v) for v in row] for row in m], dtype=complex) for m in mats]


def model_from_input(lib: Library, cfg: dict, wk, gk) -> SpinBoson:
    nmax = int(cfg["maximumdimension"])
    sizes, caps = driver_blocks(lib, wk, ecut=3000.0, high_frequency_bound=float(cfg.get("highfrequencybound", 10.0)),
                                target_dimension=int(cfg["targetdimension"]), maximum_dimension=nmax,
                                maximum_dimension_high_frequency=int(cfg.get("maximumdimensionhighfrequency", nmax)))
    return SpinBoson(wk=wk, gk=gk, block_sizes=sizes, block_nmax=caps, nspf=int(cfg["nspf"]),
                     nspf_lower=int(cfg["nspflower"]), seed=int(cfg.get("seed", 0)),
                     hsys=np.array([[_cplx(v) for v in row] for row in cfg["hsys"]], dtype=complex),
                     scoup=np.array([[_cplx(v) for v in row] for row in cfg["scoup"]], dtype=complex),
                     psi0=np.array([_cplx(v) for v in cfg["psi0"]], dtype=complex))


def _g(x: float) -> str:
    return format(float(x), ".16g")


def run(lib: Library, cfg: dict, wk, gk, out: TextIO, max_steps: Optional[int] = None, log: TextIO = sys.stderr):
    """Build the run described by `cfg` over `lib` and write its output rows to `out`."""
    model = model_from_input(lib, cfg, wk, gk)
    if int(cfg.get("nspfcapacity", model.nspf)) > model.nspf or \
            int(cfg.get("nspflowercapacity", model.nspf_lower)) > model.nspf_lower:
        print("bond growth (subspace expansion) is not available: bonds stay at their initial dimensions",
              file=log)
    tree, sop = sb_problem(lib, model)
    # defaults of the reference driver (src/spin_boson.cpp:251-255)
    ig = Integrator(lib, tree, sop, dt=float(cfg["dt"]), krylov_dim=int(cfg.get("krylovdimension", 6)),
                    krylov_tol=float(cfg.get("krylovtolerance", 1e-12)))
    try:
        ig.set_sb_initial_state(model)
        ops = observables(cfg)
        hranks = " ".join(str(r) for _, r in ig.shapes[1:]) + " " if cfg.get("printbonddimension") else ""
        nnodes = len(ig.shapes)
        out.write("0 " + "".join(_g(ig.expect(0, o).real) + " " for o in ops) + " 0 0 0 " + hranks + "\n")
        tmax, steps = float(cfg["tmax"]), 0
        while ig.time() < tmax and (max_steps is None or steps < max_steps):
            ig.step(1)
            steps += 1
            vals = "".join(_g(ig.expect(0, o).real) + " " for o in ops)
            nh = ig.counters()[0]
            out.write(f"{_g(ig.time())} {vals}0 0 {_g(nh / nnodes)} {hranks}\n")
            out.flush()
        out.write(f"nh_app: {_g(ig.counters()[0] / nnodes)}\n")
    finally:
        ig.destroy()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("input")
    ap.add_argument("--bath", required=True, help="w_k g_k pairs, or the reference's stderr log of them")
    ap.add_argument("--output", help="output file (default: the input's 'output file', else stdout)")
    ap.add_argument("--steps", type=int, help="stop after this many steps")
    args = ap.parse_args(argv)
    from . import load_library

    cfg = load_input(args.input)
    wk, gk = load_bath(args.bath)
    if "nmodes" in cfg and int(cfg["nmodes"]) != len(wk):
        raise SystemExit(f"the input asks for {cfg['nmodes']} modes, the bath file has {len(wk)}")
    lib = load_library()
    path = args.output or cfg.get("outputfile")
    if path:
        with open(path, "w") as f:
            run(lib, cfg, wk, gk, f, args.steps)
    else:
        run(lib, cfg, wk, gk, sys.stdout, args.steps)
    return 0


if __name__ == "__main__":
    sys.exit(main())
