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

Several inputs (each with its own --bath, in order) run as one batch: all their instances
advance together in one persistent launch per step (ttn_batch_step), in place of the
reference's process pool over inputs (run_examples.py); each writes its own output file.

usage: python -m ttns_lib_amd.driver input.json [input2.json ...] --bath bath.txt [--bath bath2.txt ...]
                                     [--output out] [--steps N]
"""
from __future__ import annotations

import argparse
import json
import re
import sys
from typing import Optional, TextIO

import numpy as np

from .capi import Batch, Integrator, Library, SpinBoson, driver_blocks, sb_problem


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


class Job:
    """One input: its integrator and the rows it writes."""

    def __init__(self, lib: Library, cfg: dict, wk, gk, out: TextIO, log: TextIO = sys.stderr):
        model = model_from_input(lib, cfg, wk, gk)
        if int(cfg.get("nspfcapacity", model.nspf)) > model.nspf or \
                int(cfg.get("nspflowercapacity", model.nspf_lower)) > model.nspf_lower:
            print("bond growth (subspace expansion) is not available: bonds stay at their initial dimensions",
                  file=log)
        self.tree, self.sop = sb_problem(lib, model)
        # defaults of the reference driver (src/spin_boson.cpp:251-255)
        self.ig = Integrator(lib, self.tree, self.sop, dt=float(cfg["dt"]),
                             krylov_dim=int(cfg.get("krylovdimension", 6)),
                             krylov_tol=float(cfg.get("krylovtolerance", 1e-12)))
        self.ig.set_sb_initial_state(model)
        self.ops = observables(cfg)
        self.hranks = " ".join(str(r) for _, r in self.ig.shapes[1:]) + " " if cfg.get("printbonddimension") else ""
        self.nnodes = len(self.ig.shapes)
        self.tmax, self.steps, self.out = float(cfg["tmax"]), 0, out
        out.write("0 " + "".join(_g(self.ig.expect(0, o).real) + " " for o in self.ops) + " 0 0 0 " +
                  self.hranks + "\n")

    def running(self, max_steps: Optional[int]) -> bool:
        return self.ig.time() < self.tmax and (max_steps is None or self.steps < max_steps)

    def record(self):
        """The row of the step just taken."""
        self.steps += 1
        vals = "".join(_g(self.ig.expect(0, o).real) + " " for o in self.ops)
        nh = self.ig.counters()[0]
        self.out.write(f"{_g(self.ig.time())} {vals}0 0 {_g(nh / self.nnodes)} {self.hranks}\n")
        self.out.flush()

    def finish(self):
        self.out.write(f"nh_app: {_g(self.ig.counters()[0] / self.nnodes)}\n")
        self.ig.destroy()


def run(lib: Library, cfg: dict, wk, gk, out: TextIO, max_steps: Optional[int] = None, log: TextIO = sys.stderr):
    """Build the run described by `cfg` over `lib` and write its output rows to `out`."""
    job = Job(lib, cfg, wk, gk, out, log)
    try:
        while job.running(max_steps):
            job.ig.step(1)
            job.record()
    finally:
        job.finish()


def run_batch(lib: Library, jobs_in, max_steps: Optional[int] = None, log: TextIO = sys.stderr):
    """Several runs, [(cfg, wk, gk, out), ...], advanced together: one batched launch per step
    over the runs that have not reached their tmax (the batch is rebuilt when one finishes).
    A batched launch advances instances that share dt, so runs are grouped by dt and the
    groups run one after another."""
    jobs = [Job(lib, cfg, wk, gk, out, log) for cfg, wk, gk, out in jobs_in]
    groups = {}
    for j in jobs:
        groups.setdefault(j.ig.params.dt, []).append(j)
    try:
        for group in groups.values():
            active = [j for j in group if j.running(max_steps)]
            while active:
                batch = Batch(lib, [j.ig for j in active])
                try:
                    while all(j.running(max_steps) for j in active):
                        batch.step(1)
                        for j in active:
                            j.record()
                finally:
                    batch.destroy()
                active = [j for j in active if j.running(max_steps)]
    finally:
        for j in jobs:
            j.finish()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("inputs", nargs="+")
    ap.add_argument("--bath", action="append", required=True,
                    help="w_k g_k pairs, or the reference's stderr log of them (one per input, in order)")
    ap.add_argument("--output", help="output file of a single input (default: its 'output file', else stdout)")
    ap.add_argument("--steps", type=int, help="stop after this many steps")
    args = ap.parse_args(argv)
    from . import load_library

    if len(args.bath) != len(args.inputs):
        raise SystemExit("give one --bath per input")
    cfgs = [load_input(p) for p in args.inputs]
    baths = [load_bath(p) for p in args.bath]
    for p, cfg, (wk, _) in zip(args.inputs, cfgs, baths):
        if "nmodes" in cfg and int(cfg["nmodes"]) != len(wk):
            raise SystemExit(f"{p}: the input asks for {cfg['nmodes']} modes, the bath file has {len(wk)}")
    lib = load_library()
    if len(cfgs) == 1:
        path = args.output or cfgs[0].get("outputfile")
        wk, gk = baths[0]
        if path:
            with open(path, "w") as f:
                run(lib, cfgs[0], wk, gk, f, args.steps)
        else:
            run(lib, cfgs[0], wk, gk, sys.stdout, args.steps)
        return 0
    if args.output or any("outputfile" not in c for c in cfgs):
        raise SystemExit("several inputs: each needs its own 'output file' key (and no --output)")
    files = [open(c["outputfile"], "w") for c in cfgs]
    try:
        run_batch(lib, [(c, wk, gk, f) for c, (wk, gk), f in zip(cfgs, baths, files)], args.steps)
    finally:
        for f in files:
            f.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
