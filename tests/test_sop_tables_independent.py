"""The compiled term tables (ttns_lib_amd/csrc/host/sop.cpp, shared by the oracle and the HIP
library) checked without them, at full size.

At a node of a state gauged towards it (every other node an isometry facing it), the
effective Hamiltonian the tables drive (coefficient_evolution_{leaf,branch},
projector_splitting_evolution_functions.hpp:65-258, with the spf and mf environments of
single_particle_operator.hpp:26-103 and mean_field_operator.hpp:37-184) must equal the sum
over the Hamiltonian's product terms of each term's projection, computed here term by term
from the operator definitions (H = Hsys + sum_b H0_b + Scoup x Hc_b, as
helpers.full_hamiltonian writes it) with plain tree contractions and no compiled tables:
child-subtree transfer matrices <phi|O|phi> bottom-up and the environment of the node
top-down.  Run on the plain small configuration's tree (16 modes), the adjoint
configuration's tree (64 modes) and a tree with two-mode leaves, at the root, a branch
node deep in the bath, the system leaf and a bath leaf.  A dropped coupling term or a
transposed environment is planted in the sum and fails the comparison."""
import numpy as np
import pytest

from helpers import make_integrator, sb_model, tree_children
from test_oracle_pinning import _call, _gauge_towards, heff_node
from ttns_lib_amd.capi import SpinBoson, mode_block
from ttns_lib_amd.workloads import c0_instance, c4_instance


def leaf_operators(lib, model):
    """{mode: (H0, Hc)} dense leaf operators of every bath leaf, from the mode definitions"""
    ops = {}
    n = len(model.wk)
    if model.block_sizes is None:
        for k in range(n):
            d = model.boson_dim[k]
            a = np.diag(np.sqrt(np.arange(1, d)), 1)
            ops[k + 1] = (np.diag(np.arange(d) * model.wk[k]).astype(complex), model.gk[k] * (a + a.T))
        return ops
    at = 0
    for b, (size, nmax) in enumerate(zip(model.block_sizes, model.block_nmax)):
        w, g = np.array(model.wk[at:at + size]), np.array(model.gk[at:at + size])
        blk = mode_block(lib, w, g, model.ecut, nmax)
        occ = blk["occupations"]
        ns = occ.shape[0]
        H0 = np.diag(occ @ w).astype(complex)
        Hc = np.zeros((ns, ns), complex)
        index = {tuple(o): s for s, o in enumerate(occ)}
        for s, o in enumerate(occ):
            for k in range(size):
                if o[k] > 0:  # lowering: <o - e_k| a_k |o> = sqrt(o_k)
                    t = tuple(o - np.eye(size, dtype=np.int64)[k])
                    Hc[s, index[t]] += g[k] * np.sqrt(o[k])
                    Hc[index[t], s] += g[k] * np.sqrt(o[k])
        ops[b + 1] = (H0, Hc)
        at += size
    return ops


def terms(lib, model):
    """[(coefficient, {mode: operator})] of H"""
    out = [(1.0, {0: np.asarray(model.hsys, complex)})]
    for m, (H0, Hc) in leaf_operators(lib, model).items():
        out.append((1.0, {m: H0}))
        out.append((1.0, {0: np.asarray(model.scoup, complex), m: Hc}))
    return out


class Tree:
    def __init__(self, integ, tree):
        self.integ = integ
        self.info, self.kids = tree_children(tree)
        self.parent = {c: p for p, cs in self.kids.items() for c in cs}
        self.A = {i: integ.get_node(i) for i in range(len(self.info))}
        # leaf modes below every node, and transfer matrices cached by the operators inside
        self.modes = {}
        for i in reversed(range(len(self.info))):
            self.modes[i] = ({self.info[i]["leaf_index"]} if not self.kids[i]
                             else set().union(*(self.modes[c] for c in self.kids[i])))
        self.cache = {}

    def slot_apply(self, i, T, M, k):
        """contract matrix M (d x d) into child slot k of node i's tensor T (rows x cols)"""
        dims = [self.integ.shapes[c][1] for c in self.kids[i]] + [T.shape[1]]
        X = T.reshape(dims)
        X = np.moveaxis(np.tensordot(M, X, axes=([1], [k])), 0, k)
        return X.reshape(T.shape)

    def transfer(self, i, ops):
        """<phi_a| O |phi_b> of subtree i (its hrank x hrank), O the term's operators inside"""
        key = (i, tuple(sorted((m, id(o)) for m, o in ops.items() if m in self.modes[i])))
        if key not in self.cache:
            self.cache[key] = self._transfer(i, ops)
        return self.cache[key]

    def _transfer(self, i, ops):
        A = self.A[i]
        if not self.kids[i]:
            O = ops.get(self.info[i]["leaf_index"])
            return A.conj().T @ (A if O is None else O @ A)
        Y = A
        for k, c in enumerate(self.kids[i]):
            Y = self.slot_apply(i, Y, self.transfer(c, ops), k)
        return A.conj().T @ Y

    def environment(self, n, ops):
        """<env_a| O |env_b> of node n from above (its hrank x hrank)"""
        if n not in self.parent:
            r = self.A[n].shape[1]
            return np.eye(r, dtype=complex)
        p = self.parent[n]
        Ep = self.environment(p, ops)
        A = self.A[p]
        Y = A
        for k, c in enumerate(self.kids[p]):
            if c != n:
                Y = self.slot_apply(p, Y, self.transfer(c, ops), k)
        Y = Y @ Ep.T
        k = self.kids[p].index(n)
        dims = [self.integ.shapes[c][1] for c in self.kids[p]] + [A.shape[1]]
        Ab = np.moveaxis(A.reshape(dims), k, 0).reshape(dims[k], -1)
        Yb = np.moveaxis(Y.reshape(dims), k, 0).reshape(dims[k], -1)
        return Ab.conj() @ Yb.T

    def action(self, n, ops, v, transposed_env=False):
        """the term's projection applied to node n's tensor v (with the environment transposed:
        a planted mistake)"""
        Y = v
        if self.kids[n]:
            for k, c in enumerate(self.kids[n]):
                Y = self.slot_apply(n, Y, self.transfer(c, ops), k)
        else:
            O = ops.get(self.info[n]["leaf_index"])
            if O is not None:
                Y = O @ Y
        E = self.environment(n, ops)
        return Y @ (E if transposed_env else E.T)


def targets(tree):
    info, kids = tree_children(tree)
    n = len(info)
    branch = [i for i in range(n) if info[i]["nchild"] > 0]
    leaves = [i for i in range(n) if info[i]["nchild"] == 0]
    sys_leaf = [i for i in leaves if info[i]["leaf_index"] == 0][0]
    return {"root": 0, "deep branch": branch[-1], "system leaf": sys_leaf, "bath leaf": leaves[len(leaves) // 2]}


def blocks_model():
    m = sb_model(12, 3, 6)
    return SpinBoson(wk=m.wk, gk=m.gk, block_sizes=[2] * 6, block_nmax=[3] * 6, nspf=6, nspf_lower=6)


@pytest.mark.parametrize("name", ["c0", "c4", "blocks"])
def test_compiled_tables_match_term_by_term_projection(orc, name):
    model = {"c0": lambda: c0_instance().model, "c4": lambda: c4_instance().model, "blocks": blocks_model}[name]()
    tree, _, integ = make_integrator(orc, model)
    tl = terms(orc, model)
    rng = np.random.default_rng(3)
    for label, n in targets(tree).items():
        _, _, parent, _ = _gauge_towards(rng, integ, tree, n)
        N = len(integ.shapes)
        for i in range(N - 1, 0, -1):
            _call(orc, "update_spf", integ.h, i)
        chain = []
        x = n
        while x in parent:
            chain.append(x)
            x = parent[x]
        for c in reversed(chain):
            _call(orc, "update_mf", integ.h, c)
        T = Tree(integ, tree)
        rows, cols = integ.shapes[n]
        v = rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))
        got = heff_node(orc, integ, n, v)
        acts = [c * T.action(n, ops, v) for c, ops in tl]
        want = sum(acts)
        tol = 1e-11 * np.abs(want).max()
        assert np.abs(got - want).max() < tol, (name, label)
        # planted mistakes: the sum without the largest coupling term, or with that term's
        # environment transposed, misses by far more than the tolerance
        k = max(range(2, len(tl), 2), key=lambda q: np.abs(acts[q]).max())
        assert np.abs(got - (want - acts[k])).max() > 100 * tol
        if n != 0:  # the root's environment is the 1 x 1 identity
            c, ops = tl[k]
            bad = c * T.action(n, ops, v, transposed_env=True)
            assert np.abs(got - (want - acts[k] + bad)).max() > 100 * tol, (name, label)
