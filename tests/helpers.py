"""Test helpers: model factories, dense brute-force references and TTNS contraction."""
import numpy as np
import scipy.linalg as sla

from ttns_lib_amd.capi import SIGMA, Integrator, SpinBoson, sb_problem


def ohmic_modes(n, alpha=0.1, wc=5.0, wmax_factor=4.0):
    """Equally spaced discretisation of J(w) = (pi/2) alpha w exp(-w/wc), g_k = sqrt(J dw / pi).

    A simple fixed rule so the (w_k, g_k) arrays are deterministic test inputs; the
    reference's orthogonal-polynomial discretisation only changes these numbers."""
    w = wc * np.linspace(wmax_factor / n, wmax_factor, n)
    dw = w[1] - w[0] if n > 1 else w[0]
    J = 0.5 * np.pi * alpha * w * np.exp(-w / wc)
    return w, np.sqrt(J * dw / np.pi)


def sb_model(n, d, r, r_lower=None, alpha=0.1, wc=5.0, seed=0, **kw):
    w, g = ohmic_modes(n, alpha, wc)
    return SpinBoson(wk=w, gk=g, boson_dim=[d] * n, nspf=r, nspf_lower=r if r_lower is None else r_lower,
                     seed=seed, **kw)


def make_integrator(lib, model, dt=0.05, krylov_dim=12, krylov_tol=1e-12, **kw):
    tree, sop = sb_problem(lib, model)
    integ = Integrator(lib, tree, sop, dt=dt, krylov_dim=krylov_dim, krylov_tol=krylov_tol, **kw)
    integ.set_sb_initial_state(model)
    return tree, sop, integ


def tree_children(tree):
    n, _ = tree.size()
    info = [tree.node(i) for i in range(n)]
    kids = {i: [] for i in range(n)}
    stack = []
    for i in range(n):
        while stack and len(kids[stack[-1]]) == info[stack[-1]]["nchild"]:
            stack.pop()
        if stack:
            kids[stack[-1]].append(i)
        if info[i]["nchild"] > 0:
            stack.append(i)
    return info, kids


def ttns_vector(integ, tree):
    """Contract a TTNS into the full state vector, modes in leaf pre-order."""
    info, kids = tree_children(tree)

    def rec(i):
        A = integ.get_node(i)
        if info[i]["nchild"] == 0:
            return A
        subs = [rec(c) for c in kids[i]]
        dims = [s.shape[1] for s in subs]
        T = A.reshape(dims + [A.shape[1]])
        for k, s in enumerate(subs):
            T = np.moveaxis(np.tensordot(s, T, axes=([1], [k])), 0, k)
        P = int(np.prod([s.shape[0] for s in subs]))
        return T.reshape(P, A.shape[1])

    return rec(0)[:, 0]


def full_hamiltonian(model):
    """Dense spin-boson Hamiltonian, built directly from its definition (no SOP compilation)."""
    n = len(model.wk)
    dims = [np.asarray(model.hsys).shape[0]] + list(model.boson_dim)

    def kron_all(ops):
        out = np.ones((1, 1), dtype=complex)
        for o in ops:
            out = np.kron(out, o)
        return out

    eye = [np.eye(x) for x in dims]
    H = kron_all([np.asarray(model.hsys)] + eye[1:])
    for k in range(n):
        dk = dims[k + 1]
        num = np.diag(np.arange(dk) * model.wk[k])
        a = np.diag(np.sqrt(np.arange(1, dk)), 1)
        ops = list(eye)
        ops[k + 1] = num
        H = H + kron_all(ops)
        ops = list(eye)
        ops[0] = np.asarray(model.scoup)
        ops[k + 1] = model.gk[k] * (a + a.T)
        H = H + kron_all(ops)
    return H


def exact_sigma(model, psi0, times, op="z"):
    H = full_hamiltonian(model)
    S = np.kron(SIGMA[op], np.eye(H.shape[0] // 2))
    w, V = sla.eigh(H)
    c = V.conj().T @ psi0
    out = []
    for t in times:
        psi = V @ (np.exp(-1j * w * t) * c)
        out.append((psi.conj() @ S @ psi).real)
    return np.array(out)


def random_hermitian(rng, n, scale=1.0):
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return scale * 0.5 * (X + X.conj().T)
