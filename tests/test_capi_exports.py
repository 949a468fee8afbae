"""The HIP library loads on a machine without a GPU and exports every entry point that
include/ttn_capi.h declares (no compute calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ttn_capi.h")
LIB = os.path.join(ROOT, "ttns_lib_amd", "_build", "libttns_amd.so")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(ttn_\w+)\s*\(", src, re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "ttns_lib_amd", "csrc")], check=True)
    return ctypes.CDLL(LIB)


def test_header_declares_the_core_entry_points():
    names = declared()
    for n in ["ttn_ctx_create", "ttn_tree_create", "ttn_sop_create", "ttn_integrator_create", "ttn_step",
              "ttn_expect", "ttn_batch_step", "ttn_sb_problem"]:
        assert n in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_call_needs_no_gpu(lib):
    a, b = ctypes.c_int(), ctypes.c_int()
    assert lib.ttn_version(ctypes.byref(a), ctypes.byref(b)) == 0
    assert (a.value, b.value) >= (0, 1)
