"""MI355X-native TDVP time step for spin-boson tree tensor networks.

The product is the HIP shared library `ttns_lib_amd/_build/libttns_amd.so` (C ABI in
include/ttn_capi.h).  `load_library()` loads it and fails loudly when it has not been
built or no GPU is visible: there is no CPU fallback.
"""
import os

from .capi import (  # noqa: F401
    SIGMA,
    Batch,
    Integrator,
    Library,
    SOPHandle,
    SpinBoson,
    Tree,
    TTNError,
    compiled_tables,
    sb_problem,
)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libttns_amd.so")

_LIB = None


def load_library(path: str = LIB_PATH) -> Library:
    """Load the HIP library (once per process) and open a context on device 0."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    _LIB = Library(path, "ttn_")
    return _LIB
