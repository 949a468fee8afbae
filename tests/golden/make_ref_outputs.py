"""Writes tests/golden/ref_outputs.json from data files the reference ships: the exact run
inputs (temp/infiles_{1,4,5}), the discretised baths those runs printed (temp/errfiles_*,
the w_k g_k lines between the "w_k \\t g_k" header and "bath discretised.") and the first rows
of their trajectories (example_outputs/*.out: t <sx> <sy> <sz> <1> Nonesite Ntwosite nh/|A|
[hranks of every non-root node in pre-order]).  Run once in a tree where the reference is
present; the tests read only the JSON.

Two more outputs (beta_10/ohmic exponential and sudden) have no run input in temp/; their
inputs are the matching files under inputs/, and their baths are the temp/errfiles_* whose
tree gives the printed initial bond dimensions and whose first step reproduces the printed
<sigma> (errfiles_2: 3.5e-12, errfiles_3: 2.2e-12; every other bath of the same tree shape
differs by 5e-9 or more).

usage: python tests/golden/make_ref_outputs.py [reference root]
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_outputs.json")
RUNS = {"superohmic_exponential": 1, "ohmic_debye": 4, "subohmic_exponential_t0": 5}
# name -> (input file under inputs/, output file, errfiles index)
MATCHED = {"ohmic_exponential": ("beta_10/ohmic/spin_boson_exponential.in",
                                 "example_outputs/beta_10/ohmic/spin_boson_exponential.out", 2),
           "ohmic_sudden": ("beta_10/ohmic/spin_boson_sudden.in", "example_outputs/beta_10/ohmic/spin_boson_sudden.out", 3)}
KEYS = ["Hsys", "Scoup", "psi0", "bath", "dt", "nspf", "nspf lower", "nspf capacity", "nspf lower capacity", "nmodes",
        "maximum dimension", "maximum dimension high frequency", "high frequency bound", "target dimension",
        "krylov tolerance", "krylov dimension", "spawning parameter", "unoccupied threshold", "minimum unoccupied",
        "seed", "print bond dimension"]
NROWS = 4


def bath(path):
    with open(path, "rb") as f:
        lines = f.read().decode().split("\n")
    start = next(i for i, ln in enumerate(lines) if ln.rstrip().endswith("g_k")) + 1
    end = next(i for i, ln in enumerate(lines) if ln.strip() == "bath discretised.")
    wg = [tuple(float(x) for x in ln.split()) for ln in lines[start:end]]
    return [w for w, _ in wg], [g for _, g in wg]


def rows(path):
    out = []
    with open(path) as f:
        for ln in f:
            parts = ln.split()
            if len(out) == NROWS or not parts or parts[0].startswith("nh_app"):
                break
            out.append(dict(values=[float(x) for x in parts[:5]], counters=[float(x) for x in parts[5:8]],
                            hranks=[int(x) for x in parts[8:]]))
    return out


def load_input(path):
    # the files under inputs/ carry a trailing comma before the closing brace
    with open(path) as f:
        return json.loads(re.sub(r",\s*}\s*$", "}", f.read()))


def main():
    fixture = {}
    for name, (inp, out, k) in MATCHED.items():
        cfg = load_input(os.path.join(REF, "inputs", inp))
        cfg = {key.replace(" ", "").lower(): v for key, v in cfg.items()}
        cfg = {key: cfg[key.replace(" ", "").lower()] for key in KEYS if key.replace(" ", "").lower() in cfg}
        wk, gk = bath(os.path.join(REF, "temp", f"errfiles_{k}"))
        fixture[name] = dict(input=cfg, output_file=out, wk=wk, gk=gk, rows=rows(os.path.join(REF, out)))
    for name, k in RUNS.items():
        with open(os.path.join(REF, "temp", f"infiles_{k}")) as f:
            cfg = json.load(f)
        wk, gk = bath(os.path.join(REF, "temp", f"errfiles_{k}"))
        assert len(wk) == cfg["nmodes"], (name, len(wk))
        fixture[name] = dict(input={key: cfg[key] for key in KEYS if key in cfg}, output_file=cfg["output file"],
                             wk=wk, gk=gk, rows=rows(os.path.join(REF, cfg["output file"])))
    with open(OUT, "w") as f:
        json.dump(fixture, f, separators=(",", ":"))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
