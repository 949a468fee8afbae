"""Shared fixtures.  The oracle (oracle/) is the checker; the HIP library is the product."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libttn_oracle.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _ensure_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return ORACLE_SO


@pytest.fixture(scope="session")
def orc():
    from ttns_lib_amd.capi import Library

    return Library(_ensure_oracle(), "orc_", extra=ORACLE_EXTRA)


@pytest.fixture(scope="session")
def hip():
    """The HIP product library on cuda:0; fails loudly if it is not built."""
    from ttns_lib_amd import load_library

    return load_library()


import ctypes as _C  # noqa: E402

_d = _C.POINTER(_C.c_double)
_i = _C.POINTER(_C.c_int64)
ORACLE_EXTRA = {
    "heff_node": [_C.c_void_p, _C.c_int64, _d, _d],
    "heff_bond": [_C.c_void_p, _C.c_int64, _d, _d],
    "update_spf": [_C.c_void_p, _C.c_int64],
    "update_mf": [_C.c_void_p, _C.c_int64],
    "env_get": [_C.c_void_p, _C.c_int64, _C.c_int64, _C.c_int32, _d],
    "env_set": [_C.c_void_p, _C.c_int64, _C.c_int64, _C.c_int32, _d],
    "expv": [_C.c_int64, _d, _d, _C.c_double, _C.c_double, _C.c_double, _C.c_int64, _C.c_double, _i],
    "expm": [_C.c_int64, _d, _d],
    "qr": [_C.c_int64, _C.c_int64, _d, _d, _d],
    "node_expv": [_C.c_void_p, _C.c_int64, _C.c_double, _C.c_double, _C.c_double, _C.c_int64, _C.c_double, _i],
    "normals_used": [_C.c_void_p, _i],
    "singular_values": [_C.c_int64, _d, _d],
    "tridiag_eigen": [_C.c_int64, _d, _d, _d, _d],
    "sb_adjoint_run": [_C.c_void_p, _C.c_void_p, _C.c_int64, _C.c_int64, _C.c_int64, _C.c_int64, _d, _d, _d, _d],
}
