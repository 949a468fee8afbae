"""Spin-boson driver over the HIP library, a drop-in for the reference's `mlmctdh.x input.json`
(src/spin_boson.cpp:182-700):

* reads the reference's JSON input.  Keys are matched with spaces removed, as its loader does
  (e.g. "krylov tolerance" -> krylovtolerance, :253).
* discretises the input's spectral density into nmodes harmonic modes (:292-365; the Gauss rule
  of ttns_lib_amd.bath), or takes the pairs w_k g_k from `--bath`: a two-column file, or the log
  that the reference writes to stderr, whose pairs sit between the "w_k \\t g_k" header and the
  line "bath discretised." (:356-363).
* builds the leaves with partition_modes and the harmonic mode blocks (:418-441, 515-538), the
  tree, the Hamiltonian and the seeded initial state.
* steps while t < tmax and writes the reference's output rows (`run`, :41-107):
  `t <O_1> ... <O_n> Nonesite Ntwosite nh/|A| [hranks of the non-root nodes]`, 16 significant
  digits, with the t = 0 row first.

* runs the reference driver's integrator, the subspace-expansion one (:650-657), with its
  defaults: spawning parameter and unoccupied threshold 1e-12, minimum unoccupied = the nspf
  capacity, capacities at least the initial bond dimensions (:257-278).  The bonds grow up to
  the capacities; Nonesite / Ntwosite and the current bond dimensions go into the rows.

The device holds bond dimensions up to 64.  An input whose capacities exceed that (the
reference's examples ask for 192) is refused unless --max-capacity lowers them; with it, a
note on stderr marks the first step at which a bond reaches a lowered capacity, from where the
run may differ from the reference's.

Several inputs (each with its own --bath, in order) advance together, in place of the
reference's process pool over inputs (run_examples.py); each writes its own output file.  Runs
that share dt and whose bonds are all at capacity step in one persistent launch
(ttn_batch_step); runs whose bonds can still grow step one after another.

usage: python -m ttns_lib_amd.driver input.json [input2.json ...] [--bath bath.txt [--bath bath2.txt ...]]
                                     [--output out] [--steps N] [--max-capacity C]
"""
from __future__ import annotations

import argparse
import json
import re
import sys
from typing import Optional, TextIO

import numpy as np

from .capi import MODE_SUBSPACE_EXPANSION, Batch, Integrator, Library, SpinBoson, driver_blocks, sb_problem


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
        return load_bath_text(f.read().decode(), path)


def load_bath_text(text: str, what: str = "bath"):
    """(wk, gk) from the text of a two-column file or of the reference's stderr log."""
    lines = text.split("\n")
    head = [i for i, ln in enumerate(lines) if ln.rstrip().endswith("g_k")]
    if head:
        end = next((i for i, ln in enumerate(lines) if ln.strip() == "bath discretised."), len(lines))
        lines = lines[head[0] + 1:end]
    pairs = [ln.split() for ln in lines if ln.strip() and not ln.lstrip().startswith("#")]
    wk = np.array([float(p[0]) for p in pairs])
    gk = np.array([float(p[1]) for p in pairs])
    if len(wk) == 0:
        raise ValueError(f"{what}: no w_k g_k pairs")
    return wk, gk


def _cplx(x):
    return complex(x.replace(" ", "")) if isinstance(x, str) else complex(x)


def observables(cfg: dict):
    ops = cfg["ops"]
    mats = [ops] if np.ndim(np.array(ops, dtype=object)) == 2 else ops
    return [np.array([[_cplx(v) for v in row] for row in m], dtype=complex) for m in mats]


def capacities(cfg: dict):
    """(nspf capacity, nspf lower capacity): at least the initial bond dimensions (:269-275)"""
    nspf, lower = int(cfg["nspf"]), int(cfg["nspflower"])
    return max(int(cfg.get("nspfcapacity", nspf)), nspf), max(int(cfg.get("nspflowercapacity", lower)), lower)


def model_from_input(lib: Library, cfg: dict, wk, gk, max_capacity: Optional[int] = None) -> SpinBoson:
    nmax = int(cfg["maximumdimension"])
    cap, lower_cap = capacities(cfg)
    if max_capacity is not None:
        cap, lower_cap = min(cap, max(max_capacity, int(cfg["nspf"]))), min(lower_cap, max(max_capacity,
                                                                                          int(cfg["nspflower"])))
    sizes, caps = driver_blocks(lib, wk, ecut=3000.0, high_frequency_bound=float(cfg.get("highfrequencybound", 10.0)),
                                target_dimension=int(cfg["targetdimension"]), maximum_dimension=nmax,
                                maximum_dimension_high_frequency=int(cfg.get("maximumdimensionhighfrequency", nmax)))
    return SpinBoson(wk=wk, gk=gk, block_sizes=sizes, block_nmax=caps, nspf=int(cfg["nspf"]),
                     nspf_lower=int(cfg["nspflower"]), nspf_cap=cap, nspf_lower_cap=lower_cap, seed=int(cfg.get("seed", 0)),
                     hsys=np.array([[_cplx(v) for v in row] for row in cfg["hsys"]], dtype=complex),
                     scoup=np.array([[_cplx(v) for v in row] for row in cfg["scoup"]], dtype=complex),
                     psi0=np.array([_cplx(v) for v in cfg["psi0"]], dtype=complex))


def _g(x: float) -> str:
    return format(float(x), ".16g")


class Job:
    """One input: its integrator and the rows it writes."""

    def __init__(self, lib: Library, cfg: dict, wk, gk, out: TextIO, log: TextIO = sys.stderr,
                 max_capacity: Optional[int] = None):
        model = model_from_input(lib, cfg, wk, gk, max_capacity)
        self.lowered = (model.nspf_cap, model.nspf_lower_cap) != capacities(cfg)
        if self.lowered:
            print(f"capacities lowered from {capacities(cfg)} to {(model.nspf_cap, model.nspf_lower_cap)}", file=log)
        self.log = log
        self.tree, self.sop = sb_problem(lib, model)
        cap = model.nspf_cap
        # defaults of the reference driver (src/spin_boson.cpp:251-278)
        self.ig = Integrator(lib, self.tree, self.sop, dt=float(cfg["dt"]),
                             krylov_dim=int(cfg.get("krylovdimension", 6)),
                             krylov_tol=float(cfg.get("krylovtolerance", 1e-12)), mode=MODE_SUBSPACE_EXPANSION,
                             spawning_threshold=float(cfg.get("spawningparameter", 1e-12)),
                             unoccupied_threshold=float(cfg.get("unoccupiedthreshold", 1e-12)),
                             minimum_unoccupied=int(cfg.get("minimumunoccupied", capacities(cfg)[0])))
        self.ig.set_sb_initial_state(model)
        self.caps = self.tree.capacities()
        self.ops = observables(cfg)
        # every observable is recorded on the device after each step when they fit its 8 slots
        self.on_device = 0 < len(self.ops) <= 8 and lib.has("set_observables")
        if self.on_device:
            self.ig.set_observables([(0, o) for o in self.ops])
        self.print_hranks = bool(cfg.get("printbonddimension"))
        self.nnodes = len(self.ig.shapes)
        self.tmax, self.steps, self.out = float(cfg["tmax"]), 0, out
        self.noted = False
        out.write("0 " + "".join(_g(self.ig.expect(0, o).real) + " " for o in self.ops) + " 0 0 0 " +
                  self.hranks() + "\n")

    def hranks(self) -> str:
        return " ".join(str(r) for r in self.ig.hranks()[1:]) + " " if self.print_hranks else ""

    def can_grow(self) -> bool:
        return bool(np.any(np.asarray(self.ig.hranks()) < self.caps))

    def running(self, max_steps: Optional[int]) -> bool:
        return self.ig.time() < self.tmax and (max_steps is None or self.steps < max_steps)

    def record(self):
        """The row of the step just taken."""
        self.steps += 1
        if self.on_device:
            vals = "".join(_g(v.real) + " " for v in self.ig.observables()[0][-1])
        else:
            vals = "".join(_g(self.ig.expect(0, o).real) + " " for o in self.ops)
        nh = self.ig.counters()[0]
        n1, n2 = self.ig.expansion_counters()
        self.out.write(f"{_g(self.ig.time())} {vals}{n1} {n2} {_g(nh / self.nnodes)} {self.hranks()}\n")
        self.out.flush()
        if self.lowered and not self.noted and np.any(np.asarray(self.ig.hranks()) >= self.caps):
            self.noted = True
            print(f"step {self.steps}: a bond reached its lowered capacity; from here the run may differ from "
                  "one with the input's capacities", file=self.log)

    def finish(self):
        self.out.write(f"nh_app: {_g(self.ig.counters()[0] / self.nnodes)}\n")
        self.ig.destroy()


def run(lib: Library, cfg: dict, wk, gk, out: TextIO, max_steps: Optional[int] = None, log: TextIO = sys.stderr,
        max_capacity: Optional[int] = None):
    """Build the run described by `cfg` over `lib` and write its output rows to `out`."""
    job = Job(lib, cfg, wk, gk, out, log, max_capacity)
    try:
        while job.running(max_steps):
            job.ig.step(1)
            job.record()
    finally:
        job.finish()


def run_batch(lib: Library, jobs_in, max_steps: Optional[int] = None, log: TextIO = sys.stderr,
              max_capacity: Optional[int] = None):
    """Several runs, [(cfg, wk, gk, out), ...], advanced together step by step.  A batched
    launch advances instances that share dt and cannot grow, so runs are grouped by dt; within
    a group the runs whose bonds can still grow step one at a time, the others in one batched
    launch (rebuilt whenever the set changes)."""
    jobs = [Job(lib, cfg, wk, gk, out, log, max_capacity) for cfg, wk, gk, out in jobs_in]
    groups = {}
    for j in jobs:
        groups.setdefault(j.ig.params.dt, []).append(j)
    try:
        for group in groups.values():
            active = [j for j in group if j.running(max_steps)]
            while active:
                fixed = [j for j in active if not j.can_grow()]
                for j in active:
                    if j not in fixed:
                        j.ig.step(1)
                        j.record()
                if fixed:
                    batch = Batch(lib, [j.ig for j in fixed])
                    try:
                        batch.step(1)
                        for j in fixed:
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
    ap.add_argument("--bath", action="append",
                    help="w_k g_k pairs, or the reference's stderr log of them (one per input, in order); "
                         "without it the input's spectral density is discretised (ttns_lib_amd.bath)")
    ap.add_argument("--output", help="output file of a single input (default: its 'output file', else stdout)")
    ap.add_argument("--steps", type=int, help="stop after this many steps")
    ap.add_argument("--max-capacity", type=int,
                    help="lower the bond capacities to at most this (the device holds up to 64)")
    args = ap.parse_args(argv)
    from . import load_library

    cfgs = [load_input(p) for p in args.inputs]
    if args.bath:
        if len(args.bath) != len(args.inputs):
            raise SystemExit("give one --bath per input, or none")
        baths = [load_bath(p) for p in args.bath]
    else:
        from .bath import discretise_input

        baths = []
        for cfg in cfgs:
            wk, gk = discretise_input(cfg)
            # the reference prints the pairs it discretised to stderr (src/spin_boson.cpp:354-363)
            print("w_k \t g_k", file=sys.stderr)
            for w, g in zip(wk, gk):
                print(f"{float(w)!r} {float(g)!r}", file=sys.stderr)
            print("bath discretised.", file=sys.stderr)
            baths.append((wk, gk))
    for p, cfg, (wk, _) in zip(args.inputs, cfgs, baths):
        if "nmodes" in cfg and int(cfg["nmodes"]) != len(wk):
            raise SystemExit(f"{p}: the input asks for {cfg['nmodes']} modes, the bath file has {len(wk)}")
    lib = load_library()
    if len(cfgs) == 1:
        path = args.output or cfgs[0].get("outputfile")
        wk, gk = baths[0]
        if path:
            with open(path, "w") as f:
                run(lib, cfgs[0], wk, gk, f, args.steps, max_capacity=args.max_capacity)
        else:
            run(lib, cfgs[0], wk, gk, sys.stdout, args.steps, max_capacity=args.max_capacity)
        return 0
    if args.output or any("outputfile" not in c for c in cfgs):
        raise SystemExit("several inputs: each needs its own 'output file' key (and no --output)")
    files = [open(c["outputfile"], "w") for c in cfgs]
    try:
        run_batch(lib, [(c, wk, gk, f) for c, (wk, gk), f in zip(cfgs, baths, files)], args.steps,
                  max_capacity=args.max_capacity)
    finally:
        for f in files:
            f.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
