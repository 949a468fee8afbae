"""A single branch node with explicit product terms, set up in the oracle.

The node X (two leaf children a, b of dimensions d0, d1, bond r) hangs below a root whose
other child is a leaf c of dimension r.  Term t of the SOP acts with dense random Hermitian
matrices on modes a, b and c.  With the leaves set to identities and the root to I/sqrt(r),
the oracle's own environment updates give single-particle matrices O0_t, O1_t (leaves a,
b) and mean fields MF_t (node X), read back through its test entry points, so the oracle's
H-action, mean fields and Krylov propagator at X are the reference for the node-level GPU
path fed with the same explicit terms.
"""
import ctypes as C

import numpy as np

from helpers import random_hermitian
from ttns_lib_amd.capi import OP_DENSE, Integrator, SOPHandle, Tree, compiled_tables

X_NODE, LEAF_A, LEAF_B, LEAF_C = 1, 2, 3, 4


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def env_get(orc, integ, node, term, which_mf, dim):
    out = np.zeros(2 * dim * dim)
    orc._call("env_get", integ.h, node, term, which_mf, _dp(out))
    return out.view(np.complex128).reshape(dim, dim)


def parse_tables(tab):
    """Compiled term tables (ttn_sop_compiled_tables layout) -> per node a list of terms
    {id_spf, id_mf, spf: [[(nu, cri), ...], ...], mf: [(parent_term, [(nu, cri), ...]), ...]}."""
    nodes, p = [], 0
    while p < len(tab):
        ncs, ncm, nst = (int(x) for x in tab[p:p + 3])
        p += 3
        terms = []
        for _ in range(ncs + ncm + nst):
            id_spf, id_mf, nspf = (int(x) for x in tab[p:p + 3])
            p += 3
            spf = []
            for _ in range(nspf):
                ln = int(tab[p])
                p += 1
                spf.append([(int(tab[p + 2 * k]), int(tab[p + 2 * k + 1])) for k in range(ln)])
                p += 2 * ln
            nmf = int(tab[p])
            p += 1
            mf = []
            for _ in range(nmf):
                par, nsib = int(tab[p]), int(tab[p + 1])
                p += 2
                mf.append((par, [(int(tab[p + 2 * k]), int(tab[p + 2 * k + 1])) for k in range(nsib)]))
                p += 2 * nsib
            terms.append(dict(id_spf=id_spf, id_mf=id_mf, spf=spf, mf=mf))
        nodes.append(terms)
    return nodes


class NodeProblem:
    def __init__(self, orc, d0, d1, r, nterms, seed=0, scale=1.0):
        rng = np.random.default_rng(seed)
        self.orc, self.d0, self.d1, self.r, self.nt = orc, d0, d1, r, nterms
        ops = []
        for t in range(nterms):
            for mode, d in enumerate((d0, d1, r)):
                ops.append(dict(mode=mode, kind=OP_DENSE, data=random_hermitian(rng, d, scale), r=[t]))
        self.tree = Tree.create(orc, [-1, 0, 1, 1, 0], [1, r, d0, d1, r], [1, r, d0, d1, r], [0, 0, d0, d1, r])
        self.sop = SOPHandle.create(orc, [d0, d1, r], nterms, ops)
        self.integ = Integrator(orc, self.tree, self.sop, dt=0.05, krylov_dim=16)
        ig = self.integ
        ig.set_node(LEAF_A, np.eye(d0))
        ig.set_node(LEAF_B, np.eye(d1))
        ig.set_node(LEAF_C, np.eye(r))
        ig.set_node(0, (np.eye(r) / np.sqrt(r)).reshape(r * r, 1))
        self.v = rng.standard_normal((d0 * d1, r)) + 1j * rng.standard_normal((d0 * d1, r))
        ig.set_node(X_NODE, self.v)
        for n in (LEAF_C, LEAF_B, LEAF_A):
            orc._call("update_spf", ig.h, n)
        orc._call("update_mf", ig.h, X_NODE)
        tables = parse_tables(compiled_tables(orc, self.tree, self.sop))
        self.x_terms = tables[X_NODE]
        assert len(self.x_terms) == nterms and all(not t["id_spf"] and not t["id_mf"] for t in self.x_terms)
        op0, op1, mf, self.cri = [], [], [], []
        for k, t in enumerate(self.x_terms):
            assert len(t["spf"]) == 1 and sorted(nu for nu, _ in t["spf"][0]) == [0, 1]
            chain = dict(t["spf"][0])
            op0.append(env_get(orc, ig, LEAF_A, chain[0], 0, d0))
            op1.append(env_get(orc, ig, LEAF_B, chain[1], 0, d1))
            mf.append(env_get(orc, ig, X_NODE, k, 1, r))
            self.cri.append(chain)
        self.op0, self.op1, self.mf = np.array(op0), np.array(op1), np.array(mf)

    def heff(self, v):
        v = np.ascontiguousarray(v, dtype=complex)
        out = np.zeros_like(v)
        self.orc._call("heff_node", self.integ.h, X_NODE, _dp(v.view(np.float64)), _dp(out.view(np.float64)))
        return out

    def mf_child(self, child):
        """Oracle mean fields of child `child` for every X term (update_mf on that leaf)."""
        leaf, d = (LEAF_A, self.d0) if child == 0 else (LEAF_B, self.d1)
        self.integ.set_node(X_NODE, self.v)
        self.orc._call("update_mf", self.integ.h, leaf)
        return np.array([env_get(self.orc, self.integ, leaf, self.cri[k][child], 1, d) for k in range(self.nt)])

    def expv(self, v, tau, kdim, tol):
        self.integ.set_node(X_NODE, v)
        na = C.c_int64(0)
        self.orc._call("node_expv", self.integ.h, X_NODE, 0.0, -1.0, tau, kdim, tol, C.byref(na))
        return self.integ.get_node(X_NODE), na.value
