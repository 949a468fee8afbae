"""GPU: the allocation-free node H-action (ttn_node_heff_noalloc) against the allocating node
path (ttn_node_heff) on the same operators and tensor.  The caller owns every device buffer
(allocated here through the HIP runtime); the entry point works in the library's fixed scratch.  Two calls
in a row, the second on a smaller node, then a node that needs several chunks of terms, and a
node too large for the scratch, which must be refused before any work is enqueued."""
import ctypes as C

import numpy as np
import pytest

from ttns_lib_amd.capi import Node, TTNError

pytestmark = pytest.mark.gpu


def _crandn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _operands(d0, d1, r, nt, seed):
    rng = np.random.default_rng(seed)
    op0 = _crandn(rng, nt, d0, d0) / d0
    op1 = _crandn(rng, nt, d1, d1) / d1
    mf = _crandn(rng, nt, r, r) / r
    coeff = _crandn(rng, nt)
    a = _crandn(rng, d0 * d1, r)
    return coeff, op0, op1, mf, a


class _Hip:
    """Device buffers through the HIP runtime the library itself links (libamdhip64.so.7): the
    test is the caller and allocates; the entry point under test does not."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        assert rc == 0, f"{what}: hip error {rc}"

    def upload(self, x):
        x = np.ascontiguousarray(x, dtype=np.complex128)
        p = C.c_void_p()
        self.check(self.rt.hipMalloc(C.byref(p), max(x.nbytes, 16)), "hipMalloc")
        self.check(self.rt.hipMemcpy(p, x.ctypes.data, x.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, shape):
        out = np.empty(shape, dtype=np.complex128)
        self.check(self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out


def _noalloc(hip, rt, d0, d1, r, nt, coeff, op0, op1, mf, a):
    ins = [rt.upload(x) for x in (coeff, op0, op1, mf, a)]
    res = rt.upload(np.full((d0 * d1, r), complex(np.nan, np.nan)))
    try:
        rt.check(rt.rt.hipDeviceSynchronize(), "sync")
        hip._call("node_heff_noalloc", hip.ctx, d0, d1, r, nt, *ins, res, None)
        rt.check(rt.rt.hipDeviceSynchronize(), "sync")
        return rt.download(res, (d0 * d1, r))
    finally:
        for p in ins + [res]:
            rt.rt.hipFree(p)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shapes", [[(32, 24, 16, 9), (12, 10, 8, 3)], [(64, 64, 64, 40), (20, 16, 24, 5)]])
def test_noalloc_heff_matches_node_heff_twice(hip, shapes):
    rt = _Hip()
    for k, (d0, d1, r, nt) in enumerate(shapes):
        print(f"[noalloc heff] call {k}: {d0}x{d1}x{r} terms {nt}", flush=True)
        coeff, op0, op1, mf, a = _operands(d0, d1, r, nt, seed=100 + 7 * k + d0)
        ref = Node(hip, op0, op1, mf, coeff)
        ref.set(a)
        want = ref.heff()
        ref.destroy()
        got = _noalloc(hip, rt, d0, d1, r, nt, coeff, op0, op1, mf, a)
        assert np.all(np.isfinite(got))
        assert _rel(got, want) < 1e-13


def test_noalloc_heff_refuses_a_node_beyond_the_scratch(hip):
    # (256 x 256) x 64 entries are 64 MiB per term: one term's two intermediates exceed the
    # 64 MiB scratch, so the call is refused on the host before any work is enqueued
    with pytest.raises(TTNError) as e:
        hip._call("node_heff_noalloc", hip.ctx, 256, 256, 64, 1, *[C.c_void_p(16)] * 6, None)
    assert e.value.code == 2 and "scratch" in str(e.value)
