"""The driver front-end (ttns_lib_amd/driver.py) on the reference's own example runs: it reads
the reference's input keys and the bath log format, and its output rows must match the
reference's .out files: the t = 0 row, the Nonesite / Ntwosite counters and every bond
dimension exactly over the steps in which the oracle reproduces the reference's bond growth
(test_expansion.py), <O> within the bounds the oracle meets in test_ref_outputs.py.  The
nh/|A| column follows the reference's un-vendored Krylov routine, whose H-application count
runs about 6 % below this library's; it is checked within 10 %.  On the CPU the driver runs over the
oracle library (the fake backend of these tests); the GPU half runs the command line over the
HIP library."""
import io
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from test_ref_outputs import ref_runs
from ttns_lib_amd import driver


def write_case(tmp_path, name, steps):
    run = ref_runs()[name]
    cfg = dict(run["input"])
    cfg["tmax"] = cfg["dt"] * steps
    cfg["ops"] = [[[0, 1], [1, 0]], [[0, "-1.0j"], ["1.0j", 0]], [[1, 0], [0, -1]], [[1, 0], [0, 1]]]
    cfg["output file"] = str(tmp_path / "out.out")
    inp = tmp_path / "input.json"
    # a trailing comma before the closing brace, as in the reference's inputs/
    inp.write_text(json.dumps(cfg, indent=1)[:-1] + ",\n}")
    # the reference's stderr log: a CR-separated debug line, the header, the pairs, the marker
    lines = ["0 1.0\r1 2.0\rw_k \t g_k"] + [f"{w!r} {g!r}" for w, g in zip(run["wk"], run["gk"])]
    lines += ["bath discretised.", "tree topology constructed", ""]
    bath = tmp_path / "bath.log"
    bath.write_text("\n".join(lines))
    return run, str(inp), str(bath), cfg["output file"]


def parse(text):
    rows = []
    for ln in text.splitlines():
        if ln.startswith("nh_app"):
            continue
        p = ln.split()
        rows.append(dict(values=[float(x) for x in p[:5]], counters=[float(x) for x in p[5:8]],
                         hranks=[int(x) for x in p[8:]], text=ln))
    return rows


def check_rows(rows, run, steps, bounds, exact):
    """`exact` steps with the reference's counters and bond dimensions (the steps whose
    expansion decisions rounding cannot flip), then Ntwosite within 2 attempts"""
    ref = run["rows"]
    assert len(rows) == steps + 1
    assert rows[0]["text"].startswith("0 0 0 1 1  0 0 0 ")
    if ref[0].get("hranks"):
        assert rows[0]["hranks"] == ref[0]["hranks"]
    for k, (r, e, bound) in enumerate(zip(rows[1:], ref[1:1 + steps], bounds)):
        assert abs(r["values"][0] - e["values"][0]) < 1e-15
        assert np.abs(np.array(r["values"][1:4]) - np.array(e["values"][1:4])).max() < bound
        assert abs(r["values"][4] - e["values"][4]) < 1e-12
        assert abs(r["counters"][2] / e["counters"][2] - 1.0) < 0.1
        if k < exact:
            assert r["counters"][:2] == e["counters"][:2]
            if e.get("hranks"):
                assert r["hranks"] == e["hranks"]
        else:
            assert r["counters"][0] == e["counters"][0] and abs(r["counters"][1] - e["counters"][1]) <= 2


# (run, steps, <O> bound per step, steps compared exactly).  The subohmic run's second step
# holds a two-site attempt that rounding decides: three oracle runs perturbed by 1e-15 end it
# with Ntwosite 4085 or 4086 and <sigma> 1.7e-10 .. 2.3e-10 apart (the oracle is 1.8e-10 from
# the reference there)
DRIVER_CASES = [("ohmic_sudden", 2, [1e-10, 1e-10], 2), ("subohmic_exponential_t0", 2, [5e-10, 1e-9], 1)]


def test_bath_log_and_input_parsing(tmp_path):
    run, inp, bath, _ = write_case(tmp_path, "ohmic_sudden", 1)
    wk, gk = driver.load_bath(bath)
    assert np.array_equal(wk, run["wk"]) and np.array_equal(gk, run["gk"])
    cfg = driver.load_input(inp)
    assert cfg["nspflower"] == run["input"]["nspf lower"] and cfg["krylovtolerance"] == 1e-12
    plain = tmp_path / "bath.txt"
    plain.write_text("# w g\n" + "\n".join(f"{w!r} {g!r}" for w, g in zip(run["wk"], run["gk"])))
    w2, g2 = driver.load_bath(str(plain))
    assert np.array_equal(w2, wk) and np.array_equal(g2, gk)
    assert [o.shape for o in driver.observables(cfg)] == [(2, 2)] * 4
    assert driver.observables(cfg)[1][0, 1] == -1j


@pytest.mark.parametrize("name,steps,bounds,exact", DRIVER_CASES)
def test_driver_over_oracle_matches_reference_output(orc, tmp_path, name, steps, bounds, exact):
    run, inp, bath, _ = write_case(tmp_path, name, steps)
    cfg = driver.load_input(inp)
    wk, gk = driver.load_bath(bath)
    out, log = io.StringIO(), io.StringIO()
    driver.run(orc, cfg, wk, gk, out, log=log)
    check_rows(parse(out.getvalue()), run, steps, bounds, exact)
    assert out.getvalue().rstrip().splitlines()[-1].startswith("nh_app: ")


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps,bounds,exact", DRIVER_CASES)
def test_driver_command_on_gpu_matches_reference_output(tmp_path, name, steps, bounds, exact):
    run, inp, bath, out = write_case(tmp_path, name, steps)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print(f"[driver] {name}, {steps} steps", flush=True)
    # capacity 64 in place of 192: no bond grows past 12 in these steps (the oracle's growth is
    # the same with either, test_expansion.py)
    subprocess.run([sys.executable, "-m", "ttns_lib_amd.driver", inp, "--bath", bath, "--max-capacity", "64"],
                   check=True, cwd=root, timeout=300)
    with open(out) as f:
        check_rows(parse(f.read()), run, steps, bounds, exact)


@pytest.mark.gpu
def test_driver_batch_of_inputs_on_gpu_matches_reference_output(tmp_path):
    # two inputs in one command, advanced together (one at a time while their bonds can grow)
    made = []
    for name, steps, bounds, exact in DRIVER_CASES:
        d = tmp_path / name
        d.mkdir()
        made.append((write_case(d, name, steps), steps, bounds, exact))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cmd = [sys.executable, "-m", "ttns_lib_amd.driver", "--max-capacity", "64"] + [m[0][1] for m in made]
    for m in made:
        cmd += ["--bath", m[0][2]]
    print("[driver] batch of two inputs", flush=True)
    subprocess.run(cmd, check=True, cwd=root, timeout=300)
    for (run, _, _, out), steps, bounds, exact in made:
        with open(out) as f:
            check_rows(parse(f.read()), run, steps, bounds, exact)


@pytest.mark.gpu
def test_driver_discretises_the_bath_when_no_pairs_are_given(orc, tmp_path):
    # the command without --bath discretises the input's spectral density (ttns_lib_amd.bath) and
    # prints the pairs in the reference's log format; the rows equal the oracle's on those pairs
    from ttns_lib_amd.bath import discretise_input

    cfg = {"Hsys": [[0, 1], [1, 0]], "Scoup": [[1, 0], [0, -1]], "psi0": [1, 0], "nmodes": 16, "dt": 0.05,
           "tmax": 0.1, "nspf": 6, "nspf lower": 6, "maximum dimension": 8, "target dimension": 8,
           "krylov dimension": 12, "seed": 1, "bath integration error tolerance": 1e-3,
           "bath": {"type": "exponential", "wc": 5.0, "alpha": 0.2, "s": 1},
           "ops": [[[1, 0], [0, -1]]], "output file": str(tmp_path / "out.out")}
    inp = tmp_path / "input.json"
    inp.write_text(json.dumps(cfg))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("[driver] discretised bath, 16 modes", flush=True)
    r = subprocess.run([sys.executable, "-m", "ttns_lib_amd.driver", str(inp)], check=True, cwd=root, timeout=300,
                       capture_output=True, text=True)
    wk, gk = driver.load_bath_text(r.stderr)
    c = driver.load_input(str(inp))
    w0, g0 = discretise_input(c)
    assert np.array_equal(wk, w0) and np.array_equal(gk, g0)
    out = io.StringIO()
    driver.run(orc, c, w0, g0, out, log=io.StringIO())
    with open(cfg["output file"]) as f:
        got = [ln.split() for ln in f.read().splitlines() if not ln.startswith("nh_app")]
    want = [ln.split() for ln in out.getvalue().splitlines() if not ln.startswith("nh_app")]
    assert len(got) == len(want) == 3
    for a, b in zip(got, want):
        assert abs(float(a[0]) - float(b[0])) < 1e-15 and abs(float(a[1]) - float(b[1])) < 1e-10
