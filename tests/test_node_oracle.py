"""CPU checks of the node fixture used by the node-level GPU parity tests: the operators
read back from the oracle reproduce its H-action and mean fields through the explicit
product-term form the node entry points take."""
import numpy as np
import pytest

from node_fixture import NodeProblem


@pytest.mark.parametrize("d0,d1,r,nt", [(4, 3, 5, 3), (6, 6, 4, 5)])
def test_fixture_terms_reproduce_oracle_heff(orc, d0, d1, r, nt):
    P = NodeProblem(orc, d0, d1, r, nt, seed=3)
    v = P.v
    T = v.reshape(d0, d1, r)
    want = sum(np.einsum("ij,kl,jlq,pq->ikp", P.op0[t], P.op1[t], T, P.mf[t]) for t in range(nt))
    assert np.abs(P.heff(v) - want.reshape(d0 * d1, r)).max() < 1e-12 * np.abs(want).max()


def test_fixture_mean_fields_are_hermitian(orc):
    P = NodeProblem(orc, 5, 4, 3, 4, seed=1)
    for child in (0, 1):
        mf = P.mf_child(child)
        assert np.abs(mf - np.conj(np.transpose(mf, (0, 2, 1)))).max() < 1e-12
