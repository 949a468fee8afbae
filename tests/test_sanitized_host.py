"""The host code shared by the oracle and the product (tree building, SOP compilation, mode
blocks, the spin-boson builder) and the oracle itself, run under the address and
undefined-behaviour sanitizers: the C++ drop-in driver over the oracle
(oracle/_build/sb_driver_orc_san, oracle/Makefile) on fixed-step, growing, mode-block,
mid-run-change and random one-site cases must finish without a sanitizer report and print the
rows of the unsanitized build."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = [
    "8 6 4 4 0.05 0.1 5.0 0 12",
    "8 6 2 4 0.05 0.1 5.0 0 12 --capacity 6 --spawn 1e-8 --min-unoccupied 2 --dt2 0.025 --dt2-at 2",
    "16 10 8 3 0.05 0.1 5.0 0 12 3000.0 2.0 100 6",
    "8 6 4 5 0.05 0.1 5.0 1 12 --coeff0 2.0 --coeff0-at 2 --reset-at 3 --clear-at 4",
    "4 3 1 3 0.05 0.1 5.0 0 12 --capacity 9 --spawn -1 --min-unoccupied 2",
]


def _binary(name):
    path = os.path.join(ROOT, "oracle", "_build", name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), f"_build/{name}"], check=True)
    return path


@pytest.mark.parametrize("case", CASES)
def test_host_code_is_clean_under_sanitizers(case):
    san, plain = _binary("sb_driver_orc_san"), _binary("sb_driver_orc")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([san] + case.split(), capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-2000:]
    p = subprocess.run([plain] + case.split(), capture_output=True, text=True, timeout=600, check=True)
    assert r.stdout == p.stdout
