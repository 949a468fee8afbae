"""Invalid inputs end in TTN_ERR_INVALID (code 3) with a message, never in a device launch:
the host-side checks shared by the oracle and the HIP library (tree, operators), and the
device path's own limits."""
import ctypes as C

import numpy as np
import pytest

from helpers import make_integrator, sb_model
from ttns_lib_amd.capi import MODE_SUBSPACE_EXPANSION, OP_DENSE, Batch, Integrator, SOPHandle, Tree, TTNError, sb_problem

INVALID = 3


def _tree(lib, parent, hrank, cap, leaf):
    return Tree.create(lib, parent, hrank, cap, leaf)


def check_tree_errors(lib):
    cases = [
        ([0, 0, 0], [1, 2, 2], [1, 2, 2], [0, 2, 2]),          # node 0 is not the root
        ([-1, 0, 0], [1, 3, 2], [1, 3, 2], [0, 2, 2]),         # leaf hrank above its dimension
        ([-1, 0, 1], [1, 2, 2], [1, 2, 2], [0, 2, 2]),         # leaf with a child
        ([-1, 0, 0], [1, 2, 2], [1, 1, 2], [0, 2, 2]),         # capacity below the bond dimension
        ([-1, 0, 0], [2, 2, 2], [2, 2, 2], [0, 2, 2]),         # root hrank not 1
        ([-1, 2, 0], [1, 2, 2], [1, 2, 2], [0, 2, 2]),         # parent after child
    ]
    for parent, hrank, cap, leaf in cases:
        with pytest.raises(TTNError) as e:
            _tree(lib, parent, hrank, cap, leaf)
        assert e.value.code == INVALID, (parent, hrank, cap, leaf)
        assert str(e.value)


def check_operator_errors(lib):
    op = dict(mode=0, kind=OP_DENSE, data=np.eye(3), r=[0])   # mode 0 has dimension 2
    with pytest.raises(TTNError) as e:
        SOPHandle.create(lib, [2, 2], 1, [op])
    assert e.value.code == INVALID
    op = dict(mode=5, kind=OP_DENSE, data=np.eye(2), r=[0])   # no such mode
    with pytest.raises(TTNError) as e:
        SOPHandle.create(lib, [2, 2], 1, [op])
    assert e.value.code == INVALID


def check_integrator_errors(lib):
    m = sb_model(4, 4, 4)
    tree, sop = sb_problem(lib, m)
    with pytest.raises(TTNError) as e:
        Integrator(lib, tree, sop, dt=0.0)
    assert e.value.code == INVALID
    ig = Integrator(lib, tree, sop, dt=0.05)
    rows, cols = ig.shapes[1]
    buf = np.zeros(2 * (rows + 1) * cols)
    with pytest.raises(TTNError) as e:   # the C-ABI's own shape check (not the Python wrapper's)
        lib._call("state_set", ig.h, 1, buf.ctypes.data_as(C.POINTER(C.c_double)), rows + 1, cols)
    assert e.value.code == INVALID


def test_oracle_tree_errors(orc):
    check_tree_errors(orc)


def test_oracle_operator_errors(orc):
    check_operator_errors(orc)


def test_oracle_integrator_errors(orc):
    check_integrator_errors(orc)


@pytest.mark.gpu
def test_gpu_tree_operator_integrator_errors(hip):
    check_tree_errors(hip)
    check_operator_errors(hip)
    check_integrator_errors(hip)


@pytest.mark.gpu
def test_gpu_device_limits(hip):
    m = sb_model(4, 4, 4)
    tree, sop = sb_problem(hip, m)
    with pytest.raises(TTNError) as e:
        Integrator(hip, tree, sop, dt=0.05, krylov_dim=17)   # the device Krylov space holds 16
    assert e.value.code == INVALID and "krylov_dim" in str(e.value)
    # batched instances must share dt
    a = make_integrator(hip, m, dt=0.05)[2]
    b = make_integrator(hip, m, dt=0.025)[2]
    batch = Batch(hip, [a, b])
    with pytest.raises(TTNError) as e:
        batch.step(1)   # checked when the batch steps (instances may change between calls)
    assert e.value.code == INVALID and "dt" in str(e.value)
    batch.destroy()
    # an expansion instance whose bonds can still grow is stepped on its own
    g = sb_model(4, 4, 2, nspf_cap=4, nspf_lower_cap=4)
    a = make_integrator(hip, g, mode=MODE_SUBSPACE_EXPANSION)[2]
    b = make_integrator(hip, g)[2]
    batch = Batch(hip, [a, b])
    with pytest.raises(TTNError) as e:
        batch.step(1)
    assert e.value.code == INVALID and "capacity" in str(e.value)
    batch.destroy()
