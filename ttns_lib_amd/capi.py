"""ctypes bindings of the C ABI declared in include/ttn_capi.h.

`Library(path, prefix)` binds every entry point of a shared library that exports the
ttn_* call signatures under `prefix` ("ttn_" for the HIP library).  `Problem`,
`Integrator` and `Batch` are thin Python handles over those calls, used by the tests,
the benchmark and the autograd harness.  Failures raise `TTNError` with the library's
own message, mirroring the reference's exception chains
(integrators/projector_splitting_integrator.hpp:321-325).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

i64 = C.c_int64
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
dblp = C.POINTER(C.c_double)
vp = C.c_void_p

STATUS = {0: "OK", 1: "ERR_NUMERIC", 2: "ERR_CAPACITY", 3: "ERR_INVALID", 4: "ERR_DEVICE"}
OP_DENSE, OP_DIAG, OP_CSR, OP_IDENTITY = 0, 1, 2, 3
MODE_FIXED_STEP, MODE_SUBSPACE_EXPANSION = 0, 1


class TTNError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class IntegParams(C.Structure):
    _fields_ = [
        ("krylov_dim", C.c_int64),
        ("krylov_tol", C.c_double),
        ("dt", C.c_double),
        ("mode", C.c_int32),
        ("spawning_threshold", C.c_double),
        ("unoccupied_threshold", C.c_double),
        ("minimum_unoccupied", C.c_int64),
        ("neigenvalues", C.c_int64),
        ("team_size", C.c_int64),
        ("eigenvalue_tolerance", C.c_double),
        ("only_apply_when_no_unoccupied", C.c_int32),
        ("rng_seed", C.c_uint32),
    ]


class SBParamsC(C.Structure):
    _fields_ = [
        ("nmodes", C.c_int64),
        ("wk", dblp),
        ("gk", dblp),
        ("boson_dim", i64p),
        ("nspf", C.c_int64),
        ("nspf_lower", C.c_int64),
        ("nspf_cap", C.c_int64),
        ("nspf_lower_cap", C.c_int64),
        ("nhilb", C.c_int64),
        ("hsys", dblp),
        ("scoup", dblp),
        ("psi0", dblp),
        ("seed", C.c_uint64),
        ("nblocks", C.c_int64),
        ("block_sizes", i64p),
        ("block_nmax", i64p),
        ("ecut", C.c_double),
    ]


# name -> argtypes (every function returns int status unless listed in _RESTYPE)
_SIGS = {
    "ctx_create": [C.c_int, C.POINTER(vp)],
    "ctx_destroy": [vp],
    "last_error": [vp],
    "tree_create": [vp, i64, i64p, i64p, i64p, i64p, C.POINTER(vp)],
    "tree_destroy": [vp],
    "tree_node": [vp, i64, i64p, i64p, i64p, i64p],
    "tree_size": [vp, i64p, i64p],
    "tree_capacities": [vp, i64p],
    "traversal": [vp, i64p, i64p],
    "sop_create": [vp, i64, i64p, i64, dblp, i64, i32p, i32p, i64p, dblp, i64p, i64p, i64p, i64p, C.POINTER(vp)],
    "sop_destroy": [vp],
    "sop_compiled_tables": [vp, vp, i64p, i64p],
    "integrator_create": [vp, vp, vp, C.POINTER(IntegParams), C.POINTER(vp)],
    "integrator_destroy": [vp],
    "state_set": [vp, i64, dblp, i64, i64],
    "state_get": [vp, i64, dblp, i64, i64],
    "step": [vp, C.c_double, C.c_double, i64],
    "time": [vp, dblp],
    "counters": [vp, i64p, i64p],
    "expansion_counters": [vp, i64p, i64p],
    "hranks": [vp, i64p],
    "state_shape": [vp, i64, i64p, i64p],
    "set_dt": [vp, C.c_double],
    "get_dt": [vp, dblp],
    "reset_nsteps": [vp],
    "set_coefficients": [vp, i64, dblp],
    "checkpoint_save": [vp, i64p, C.c_void_p],
    "checkpoint_load": [vp, C.c_void_p, i64],
    "set_expansion_params": [vp, C.c_void_p],
    "expansion_draws": [vp, i64p],
    "set_expansion_log": [vp, i64],
    "expansion_log": [vp, i64p, dblp],
    "expect": [vp, i64, dblp, dblp],
    "norm2": [vp, dblp],
    "set_observable": [vp, i64, dblp],
    "set_observables": [vp, i64, i64p, dblp],
    "observables": [vp, i64p, dblp],
    "trajectory": [vp, i64p, dblp],
    "sb_problem": [vp, C.POINTER(SBParamsC), C.POINTER(vp), C.POINTER(vp)],
    "sb_initial_state": [vp, C.POINTER(SBParamsC)],
    "sb_partition_modes": [vp, C.c_int64, dblp, C.c_double, C.c_double, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                           i64p, i64p],
    "sb_mode_block": [vp, C.c_int64, dblp, dblp, C.c_double, C.c_int64, i64p, i64p, i64p, dblp, i64p, i64p, dblp],
    "batch_create": [vp, i64, C.POINTER(vp), C.POINTER(vp)],
    "batch_destroy": [vp],
    "batch_step": [vp, C.c_double, C.c_double, i64],
    "last_step_ms": [vp, dblp],
    "flops": [vp, dblp],
    "profile": [vp, i64p],
    "batch_last_step_ms": [vp, dblp],
    "ctx_set_tuning": [vp, i64, C.c_int],
    "ctx_set_small": [vp, i64],
    "batch_teams": [vp, i64p],
    "node_stats": [vp, i64p],
    "batch_set_graph": [vp, C.c_int],
    "batch_graph_stats": [vp, i64p, i64p],
    "node_create": [vp, i64, i64, i64, i64, dblp, dblp, dblp, dblp, C.POINTER(vp)],
    "node_destroy": [vp],
    "node_set": [vp, dblp],
    "node_get": [vp, dblp],
    "node_get_result": [vp, dblp],
    "node_heff": [vp, i64],
    "node_mf": [vp, i64, i64, dblp],
    "node_expv": [vp, C.c_double, C.c_double, C.c_double, i64, C.c_double, i64p],
    "node_heff_noalloc": [vp, i64, i64, i64, i64, vp, vp, vp, vp, vp, vp, vp],
    "node_last": [vp, dblp, dblp],
    "sb_adjoint_run": [vp, C.POINTER(SBParamsC), i64, i64, i64, i64, dblp, dblp, dblp, dblp],
    "ctx_set_autotune": [vp, C.c_int],
    "autotune_table": [vp, i64p, i64p, dblp],
    "ctx_set_heff_form": [vp, C.c_int],
    "transition_densities": [vp, vp, i64, dblp, i64, dblp],
}
_RESTYPE = {"last_error": C.c_char_p}


def _dp(a: np.ndarray):
    return a.ctypes.data_as(dblp)


def _ip(a: np.ndarray):
    return a.ctypes.data_as(i64p)


def _cplx_in(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return a.view(np.float64)


# output tiles (rows, columns per workgroup) of the node-level GEMM, in ttn_autotune_table order
GEMM_TILES = [(32, 32), (64, 32), (32, 64), (64, 64), (128, 64), (64, 128)]


class Library:
    """Binds the ttn_* signatures exported under `prefix` by the shared library `path`."""

    def __init__(self, path: str, prefix: str = "ttn_", extra: Optional[dict] = None):
        self.path = path
        self.prefix = prefix
        self.so = C.CDLL(path)
        self.fn = {}
        sigs = dict(_SIGS)
        if extra:
            sigs.update(extra)
        for name, args in sigs.items():
            sym = prefix + name
            if not hasattr(self.so, sym):
                continue
            f = getattr(self.so, sym)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
            self.fn[name] = f
        self.ctx = vp()
        self._call("ctx_create", 0, C.byref(self.ctx))

    def has(self, name: str) -> bool:
        return name in self.fn

    def set_tuning(self, team_min: int = 2048, teams: bool = True):
        """Workgroup-team knobs of the GPU library: operations on node tensors of at least
        team_min entries are spread over an instance's team; teams=False runs every
        instance on one workgroup."""
        self._call("ctx_set_tuning", self.ctx, int(team_min), int(bool(teams)))

    def set_autotune(self, enabled: bool = True):
        """tile autotuner of the node-level GEMMs (ttn_ctx_set_autotune)"""
        self._call("ctx_set_autotune", self.ctx, int(bool(enabled)))

    def set_heff_form(self, form: int = 0):
        """node H-action's mode-0 and mean-field stages: 0 autotuned, 1 staged, 2 fused
        (ttn_ctx_set_heff_form)"""
        self._call("ctx_set_heff_form", self.ctx, int(form))

    def autotune_table(self) -> list:
        """[(M, N, K, batch, kblock, tile, ksplit, ms)] of every GEMM shape the autotuner timed;
        tile indexes GEMM_TILES"""
        n = i64(0)
        self._call("autotune_table", self.ctx, C.byref(n), None, None)
        if n.value == 0:
            return []
        e = np.zeros(7 * n.value, dtype=np.int64)
        ms = np.zeros(n.value)
        self._call("autotune_table", self.ctx, C.byref(n), _ip(e), _dp(ms))
        return [tuple(int(x) for x in e[7 * i:7 * i + 7]) + (float(ms[i]),) for i in range(n.value)]

    def set_small(self, small_max: int = 512):
        """Krylov propagations of at most small_max entries run on one workgroup with the
        basis in registers (0: none)."""
        self._call("ctx_set_small", self.ctx, int(small_max))

    def error(self) -> str:
        msg = self.fn["last_error"](self.ctx)
        return msg.decode() if msg else ""

    def _call(self, name, *args):
        rc = self.fn[name](*args)
        if rc != 0:
            raise TTNError(rc, f"{self.prefix}{name}: " + (self.error() if name != "ctx_create" else ""))
        return rc

    def close(self):
        if self.ctx:
            self.fn["ctx_destroy"](self.ctx)
            self.ctx = vp()


@dataclass
class SpinBoson:
    """Spin-boson model (src/spin_boson.cpp:418-649): one bath mode per leaf with the Fock
    truncations boson_dim, or, when block_sizes is given, consecutive blocks of modes on one
    leaf each, mode k keeping min(block_nmax[b], floor(ecut / |w_k|)) states."""

    wk: Sequence[float]
    gk: Sequence[float]
    boson_dim: Optional[Sequence[int]] = None
    nspf: int = 8
    nspf_lower: int = 8
    nspf_cap: int = 0
    nspf_lower_cap: int = 0
    hsys: np.ndarray = field(default_factory=lambda: np.array([[0, 1], [1, 0]], dtype=complex))
    scoup: np.ndarray = field(default_factory=lambda: np.array([[1, 0], [0, -1]], dtype=complex))
    psi0: np.ndarray = field(default_factory=lambda: np.array([1, 0], dtype=complex))
    seed: int = 0
    block_sizes: Optional[Sequence[int]] = None
    block_nmax: Optional[Sequence[int]] = None
    ecut: float = 3000.0

    def to_c(self) -> SBParamsC:
        blocks = self.block_sizes is not None
        if not blocks and self.boson_dim is None:
            raise ValueError("SpinBoson needs boson_dim or block_sizes")
        self._keep = [
            np.ascontiguousarray(self.wk, dtype=np.float64),
            np.ascontiguousarray(self.gk, dtype=np.float64),
            np.ascontiguousarray(self.boson_dim if not blocks else [], dtype=np.int64),
            _cplx_in(self.hsys),
            _cplx_in(self.scoup),
            _cplx_in(self.psi0),
            np.ascontiguousarray(self.block_sizes if blocks else [], dtype=np.int64),
            np.ascontiguousarray(self.block_nmax if blocks else [], dtype=np.int64),
        ]
        wk, gk, bd, hs, sc, ps, bs, bn = self._keep
        nh = int(np.asarray(self.hsys).shape[0])
        return SBParamsC(
            len(wk), _dp(wk), _dp(gk), _ip(bd) if not blocks else None, self.nspf, self.nspf_lower,
            max(self.nspf_cap, self.nspf), max(self.nspf_lower_cap, self.nspf_lower), nh,
            _dp(hs), _dp(sc), _dp(ps), self.seed, len(bs) if blocks else 0,
            _ip(bs) if blocks else None, _ip(bn) if blocks else None, float(self.ecut),
        )


def partition_modes(lib: Library, wk, ecut: float, wc: float, nmax: int, nmax4: int, nmax2: int,
                    nmax3: int) -> list:
    """Block sizes of the reference's partition_modes (src/mode_data.hpp:818-886)."""
    w = np.ascontiguousarray(wk, dtype=np.float64)
    nb = i64()
    out = np.zeros(len(w), dtype=np.int64)
    lib._call("sb_partition_modes", lib.ctx, len(w), _dp(w), float(ecut), float(wc), int(nmax), int(nmax4),
              int(nmax2), int(nmax3), C.byref(nb), _ip(out))
    return [int(x) for x in out[: nb.value]]


def driver_blocks(lib: Library, wk, ecut: float = 3000.0, high_frequency_bound: float = 10.0,
                  target_dimension: int = 1000, maximum_dimension: int = 100,
                  maximum_dimension_high_frequency: Optional[int] = None):
    """(block_sizes, block_nmax) as the reference driver derives them from its input keys
    (src/spin_boson.cpp:194-209, 418-441): partition_modes with wc = 2 * high frequency bound,
    and each block's cutoff cap chosen by its first mode's frequency."""
    nhigh = maximum_dimension if maximum_dimension_high_frequency is None else maximum_dimension_high_frequency
    wc2 = 2.0 * high_frequency_bound
    sizes = partition_modes(lib, wk, ecut, wc2, target_dimension, 250, maximum_dimension, nhigh)
    nmax, at = [], 0
    for s in sizes:
        nmax.append(nhigh if wk[at] > wc2 else maximum_dimension)
        at += s
    return sizes, nmax


def mode_block(lib: Library, w, g, ecut: float, nmax: int) -> dict:
    """Occupation table, H0 diagonal and Hc (CSR) of one leaf's mode block."""
    w = np.ascontiguousarray(w, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    ns, nz = i64(), i64()
    args = (lib.ctx, len(w), _dp(w), _dp(g), float(ecut), int(nmax), C.byref(ns), C.byref(nz))
    lib._call("sb_mode_block", *args, None, None, None, None, None)
    occ = np.zeros((ns.value, len(w)), dtype=np.int64)
    h0 = np.zeros(ns.value)
    rp = np.zeros(ns.value + 1, dtype=np.int64)
    ci = np.zeros(max(nz.value, 1), dtype=np.int64)
    va = np.zeros(max(nz.value, 1))
    lib._call("sb_mode_block", *args, _ip(occ), _dp(h0), _ip(rp), _ip(ci), _dp(va))
    return dict(occupations=occ, h0=h0, rowptr=rp, colind=ci[: nz.value], vals=va[: nz.value])


class Tree:
    def __init__(self, lib: Library, handle: vp):
        self.lib, self.h = lib, handle

    @classmethod
    def create(cls, lib: Library, parent, hrank, cap_hrank, leaf_dim):
        arrs = [np.ascontiguousarray(x, dtype=np.int64) for x in (parent, hrank, cap_hrank, leaf_dim)]
        h = vp()
        lib._call("tree_create", lib.ctx, len(arrs[0]), *[_ip(a) for a in arrs], C.byref(h))
        return cls(lib, h)

    def size(self):
        n, m = i64(), i64()
        self.lib._call("tree_size", self.h, C.byref(n), C.byref(m))
        return n.value, m.value

    def node(self, i):
        a, b, c, d = i64(), i64(), i64(), i64()
        self.lib._call("tree_node", self.h, i, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return dict(nchild=a.value, rows=b.value, hrank=c.value, leaf_index=d.value)

    def capacities(self) -> np.ndarray:
        """bond capacity of every node (pre-order)"""
        n, _ = self.size()
        out = np.zeros(n, dtype=np.int64)
        self.lib._call("tree_capacities", self.h, out.ctypes.data_as(i64p))
        return out

    def shapes(self):
        n, _ = self.size()
        return [(self.node(i)["rows"], self.node(i)["hrank"]) for i in range(n)]

    def traversal(self):
        n = i64()
        self.lib._call("traversal", self.h, C.byref(n), None)
        out = np.zeros(n.value, dtype=np.int64)
        self.lib._call("traversal", self.h, C.byref(n), _ip(out))
        return out

    def destroy(self):
        if self.h:
            self.lib.fn["tree_destroy"](self.h)
            self.h = vp()


class SOPHandle:
    def __init__(self, lib: Library, handle: vp):
        self.lib, self.h = lib, handle

    @classmethod
    def create(cls, lib: Library, mode_dims, nterms, ops, coeff=None):
        """ops: list of dicts {mode, kind, data (complex array), idata (int array, CSR), r (terms)}."""
        mode_dims = np.ascontiguousarray(mode_dims, dtype=np.int64)
        coeff = np.ones(nterms, dtype=complex) if coeff is None else np.asarray(coeff, dtype=complex)
        op_mode = np.array([o["mode"] for o in ops], dtype=np.int32)
        op_kind = np.array([o["kind"] for o in ops], dtype=np.int32)
        cd, idd, rr = [], [], []
        coff, ioff, roff = [0], [0], [0]
        for o in ops:
            d = np.asarray(o.get("data", np.zeros(0)), dtype=complex).ravel()
            cd.append(d)
            coff.append(coff[-1] + d.size)
            ii = np.asarray(o.get("idata", np.zeros(0)), dtype=np.int64).ravel()
            idd.append(ii)
            ioff.append(ioff[-1] + ii.size)
            r = np.asarray(o["r"], dtype=np.int64).ravel()
            rr.append(r)
            roff.append(roff[-1] + r.size)
        cdata = _cplx_in(np.concatenate(cd) if cd else np.zeros(0, complex))
        idata = np.ascontiguousarray(np.concatenate(idd) if idd else np.zeros(0), dtype=np.int64)
        rdat = np.ascontiguousarray(np.concatenate(rr) if rr else np.zeros(0), dtype=np.int64)
        coff, ioff, roff = (np.array(x, dtype=np.int64) for x in (coff, ioff, roff))
        cf = _cplx_in(coeff)
        h = vp()
        lib._call(
            "sop_create", lib.ctx, len(mode_dims), _ip(mode_dims), nterms, _dp(cf), len(ops),
            op_mode.ctypes.data_as(i32p), op_kind.ctypes.data_as(i32p), _ip(coff), _dp(cdata),
            _ip(ioff), _ip(idata), _ip(roff), _ip(rdat), C.byref(h),
        )
        return cls(lib, h)

    def destroy(self):
        if self.h:
            self.lib.fn["sop_destroy"](self.h)
            self.h = vp()


def compiled_tables(lib: Library, tree: Tree, sop: SOPHandle) -> np.ndarray:
    n = i64(0)
    lib._call("sop_compiled_tables", tree.h, sop.h, C.byref(n), None)
    buf = np.zeros(n.value, dtype=np.int64)
    lib._call("sop_compiled_tables", tree.h, sop.h, C.byref(n), _ip(buf))
    return buf


def sb_gradient(lib: Library, model: SpinBoson, dops) -> tuple:
    """(dJ/dg_k, dJ/dw_k) from the cotangents of the spin-boson operator values
    (ttn_sb_gradient; the Fock cutoffs held fixed)"""
    d = _cplx_in(dops)
    n = len(model.wk)
    dg, dw = np.zeros(n), np.zeros(n)
    p = model.to_c()
    lib._call("sb_gradient", lib.ctx, C.byref(p), len(d) // 2, _dp(d), _dp(dg), _dp(dw))
    return dg, dw


def sb_problem(lib: Library, model: SpinBoson):
    t, s = vp(), vp()
    p = model.to_c()
    lib._call("sb_problem", lib.ctx, C.byref(p), C.byref(t), C.byref(s))
    return Tree(lib, t), SOPHandle(lib, s)


class Integrator:
    """An integrator handle (ttn_integrator_create).  mode=MODE_SUBSPACE_EXPANSION is the
    reference's subspace_expansion_projector_splitting_integrator: bonds below their capacity
    grow, so node shapes are read back from the library (`shape`, `hranks`)."""

    def __init__(self, lib: Library, tree: Tree, sop: SOPHandle, *, dt: float, krylov_dim: int = 12,
                 krylov_tol: float = 1e-12, mode: int = MODE_FIXED_STEP, team_size: int = 0,
                 spawning_threshold: float = 1e-12, unoccupied_threshold: float = 1e-12,
                 minimum_unoccupied: int = 0, neigenvalues: int = 2, eigenvalue_tolerance: float = 0.0,
                 only_apply_when_no_unoccupied: bool = False, rng_seed: int = 0):
        self.lib, self.tree = lib, tree
        self.params = IntegParams(krylov_dim, krylov_tol, dt, mode, spawning_threshold, unoccupied_threshold,
                                  minimum_unoccupied, neigenvalues, team_size, eigenvalue_tolerance,
                                  int(bool(only_apply_when_no_unoccupied)), rng_seed)
        self.h = vp()
        lib._call("integrator_create", lib.ctx, tree.h, sop.h, C.byref(self.params), C.byref(self.h))
        self.shapes = tree.shapes()

    def shape(self, i):
        """current (rows, cols) of node i"""
        if not self.lib.has("state_shape"):
            return self.shapes[i]
        r, c = i64(), i64()
        self.lib._call("state_shape", self.h, i, C.byref(r), C.byref(c))
        return r.value, c.value

    def hranks(self) -> np.ndarray:
        out = np.zeros(len(self.shapes), dtype=np.int64)
        self.lib._call("hranks", self.h, _ip(out))
        return out

    def expansion_counters(self):
        """(Nonesite, Ntwosite) of the expansion engine"""
        a, b = i64(), i64()
        self.lib._call("expansion_counters", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def set_dt(self, dt: float):
        self.lib._call("set_dt", self.h, float(dt))

    def get_dt(self) -> float:
        out = C.c_double()
        self.lib._call("get_dt", self.h, C.byref(out))
        return out.value

    def reset_nsteps(self):
        self.lib._call("reset_nsteps", self.h)

    def set_expansion_params(self, **kw):
        """changes spawning_threshold, unoccupied_threshold, minimum_unoccupied, neigenvalues,
        eigenvalue_tolerance or only_apply_when_no_unoccupied of an expansion integrator"""
        for k, v in kw.items():
            if k not in ("spawning_threshold", "unoccupied_threshold", "minimum_unoccupied", "neigenvalues",
                         "eigenvalue_tolerance", "only_apply_when_no_unoccupied"):
                raise TypeError(f"not an expansion parameter: {k}")
            setattr(self.params, k, v)
        self.lib._call("set_expansion_params", self.h, C.cast(C.pointer(self.params), C.c_void_p))

    def checkpoint(self) -> bytes:
        """the integrator's state (ttn_checkpoint_save)"""
        n = i64(0)
        self.lib._call("checkpoint_save", self.h, C.byref(n), None)
        buf = C.create_string_buffer(n.value)
        self.lib._call("checkpoint_save", self.h, C.byref(n), buf)
        return buf.raw[:n.value]

    def restore(self, data: bytes):
        """continue from a checkpoint of an integrator on the same tree, operator and parameters"""
        buf = C.create_string_buffer(bytes(data), len(data))
        self.lib._call("checkpoint_load", self.h, buf, len(data))

    def set_coefficients(self, coeff):
        """new SOP coefficients (one complex per term), used from the next step on"""
        c = np.ascontiguousarray(np.asarray(coeff, dtype=np.complex128))
        self.lib._call("set_coefficients", self.h, int(c.size), _dp(c.view(np.float64)))

    def set_expansion_log(self, capacity: int):
        self.lib._call("set_expansion_log", self.h, int(capacity))

    def expansion_log(self) -> np.ndarray:
        """records (node, direction, Ritz values, functions added, spawning values sqrt(S) dt / 2
        of the 16 largest Ritz values S)"""
        n = i64(0)
        self.lib._call("expansion_log", self.h, C.byref(n), None)
        out = np.zeros(20 * n.value)
        self.lib._call("expansion_log", self.h, C.byref(n), _dp(out))
        return out.reshape(-1, 20)

    def expansion_draws(self) -> int:
        """normal draws the expansion's trial vectors have taken so far"""
        n = i64()
        self.lib._call("expansion_draws", self.h, C.byref(n))
        return n.value

    def set_sb_initial_state(self, model: SpinBoson):
        p = model.to_c()
        self.lib._call("sb_initial_state", self.h, C.byref(p))

    def set_node(self, i, a):
        a = np.asarray(a, dtype=complex)
        rows, cols = self.shape(i)
        assert a.shape == (rows, cols), (a.shape, rows, cols)
        buf = _cplx_in(a)
        self.lib._call("state_set", self.h, i, _dp(buf), rows, cols)

    def get_node(self, i):
        rows, cols = self.shape(i)
        buf = np.zeros(2 * rows * cols, dtype=np.float64)
        self.lib._call("state_get", self.h, i, _dp(buf), rows, cols)
        return buf.view(np.complex128).reshape(rows, cols)

    def step(self, nsteps: int = 1, coeff: complex = -1j):
        self.lib._call("step", self.h, coeff.real, coeff.imag, nsteps)

    def time(self):
        t = C.c_double()
        self.lib._call("time", self.h, C.byref(t))
        return t.value

    def counters(self):
        a, b = i64(), i64()
        self.lib._call("counters", self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def expect(self, mode: int, op) -> complex:
        o = _cplx_in(op)
        out = np.zeros(2)
        self.lib._call("expect", self.h, mode, _dp(o), _dp(out))
        return complex(out[0], out[1])

    def norm2(self) -> float:
        out = C.c_double()
        self.lib._call("norm2", self.h, C.byref(out))
        return out.value

    def sb_adjoint(self, model: "SpinBoson", nsteps: int, mode: int, op, *, every: int = 20, nquad: int = 0):
        """(J, dJ/dg_k, dJ/dw_k) for J = Re<psi(T)|O|psi(T)> after nsteps real-time calls from a
        state that has not been stepped (ttn_sb_adjoint_run: costate method, the state
        checkpointed every `every` calls, nquad-point Gauss-Legendre quadrature per step, 0: chosen
        from dt and the largest frequency).  The
        integrator is left at the state after nsteps calls."""
        o = _cplx_in(op)
        val = C.c_double()
        n = len(model.wk)
        dg, dw = np.zeros(n), np.zeros(n)
        p = model.to_c()
        self.lib._call("sb_adjoint_run", self.h, C.byref(p), int(nsteps), int(every), int(nquad), int(mode), _dp(o),
                       C.byref(val), _dp(dg), _dp(dw))
        return val.value, dg, dw

    def transition_densities(self, bra: "Integrator", ins: int, S) -> list:
        """[(rho_m, rhoS_m)] per leaf mode m: rho_m(j, l) = <bra|(|j><l|)_m|self>, rhoS_m the same
        with S inserted on leaf mode ins (zero on that leaf)"""
        s = _cplx_in(S)
        leaves = [self.tree.node(i) for i in range(len(self.shapes))]
        ml = [0] * sum(1 for x in leaves if x["nchild"] == 0)
        for x in leaves:
            if x["nchild"] == 0:
                ml[x["leaf_index"]] = x["rows"]
        tot = sum(2 * d * d for d in ml)
        out = np.zeros(2 * tot)
        self.lib._call("transition_densities", self.h, bra.h, int(ins), _dp(s), tot, _dp(out))
        z = out.view(np.complex128)
        res, at = [], 0
        for d in ml:
            res.append((z[at:at + d * d].reshape(d, d).copy(), z[at + d * d:at + 2 * d * d].reshape(d, d).copy()))
            at += 2 * d * d
        return res

    def last_step_ms(self) -> float:
        out = C.c_double()
        self.lib._call("last_step_ms", self.h, C.byref(out))
        return out.value

    def flops(self) -> float:
        out = C.c_double()
        self.lib._call("flops", self.h, C.byref(out))
        return out.value

    PROF_NAMES = ["heff", "hbond", "orth", "expm", "qr", "spf", "mf", "gauge", "expv_node", "expv_bond", "obs",
                  "total", "small"]

    def profile(self) -> dict:
        """Device time per sweep phase in microseconds; "n_small" counts the propagations
        run by the small-vector propagator (their time is "small")."""
        out = np.zeros(16, dtype=np.int64)
        self.lib._call("profile", self.h, _ip(out))
        d = {k: out[i] / 100.0 for i, k in enumerate(self.PROF_NAMES)}
        d["n_small"] = int(out[len(self.PROF_NAMES)])
        return d

    def node_stats(self) -> np.ndarray:
        """per node: node H applications, their Krylov sizes summed, bond H applications, sizes,
        the device ticks of the node and of the bond propagations (100 MHz), and of the H-actions
        inside them (outside the small-vector propagator)"""
        out = np.zeros(8 * len(self.shapes), dtype=np.int64)
        self.lib._call("node_stats", self.h, _ip(out))
        return out.reshape(-1, 8)

    def set_observable(self, mode: int, op):
        self._obs = _cplx_in(op)
        self.nobs = 1
        self.lib._call("set_observable", self.h, mode, _dp(self._obs))

    def set_observables(self, obs):
        """[(mode, d x d operator), ...] (at most 8) recorded on the device after every step"""
        modes = np.array([m for m, _ in obs], dtype=np.int64)
        self._obs = np.ascontiguousarray(np.concatenate([_cplx_in(o).ravel() for _, o in obs])) if obs else \
            np.zeros(2)
        self.nobs = len(obs)
        self.lib._call("set_observables", self.h, len(obs), _ip(modes), _dp(self._obs))

    def observables(self):
        """(values (steps, nobs) complex, norms (steps,)) recorded since set_observables"""
        n = i64(0)
        self.lib._call("observables", self.h, C.byref(n), None)
        k = getattr(self, "nobs", 1)
        out = np.zeros((2 * k + 1) * n.value)
        self.lib._call("observables", self.h, C.byref(n), _dp(out))
        rec = out.reshape(-1, 2 * k + 1)
        return rec[:, 0:2 * k:2] + 1j * rec[:, 1:2 * k:2], rec[:, 2 * k]

    def trajectory(self) -> np.ndarray:
        n = i64(0)
        self.lib._call("trajectory", self.h, C.byref(n), None)
        out = np.zeros(2 * n.value)
        self.lib._call("trajectory", self.h, C.byref(n), _dp(out))
        return out.reshape(-1, 2)

    def destroy(self):
        if self.h:
            self.lib.fn["integrator_destroy"](self.h)
            self.h = vp()


class Batch:
    def __init__(self, lib: Library, integs: Sequence[Integrator]):
        self.lib = lib
        self.integs = list(integs)
        arr = (vp * len(integs))(*[i.h for i in integs])
        self.h = vp()
        lib._call("batch_create", lib.ctx, len(integs), arr, C.byref(self.h))

    def step(self, nsteps: int = 1, coeff: complex = -1j):
        self.lib._call("batch_step", self.h, coeff.real, coeff.imag, nsteps)

    def last_step_ms(self) -> float:
        out = C.c_double()
        self.lib._call("batch_last_step_ms", self.h, C.byref(out))
        return out.value

    def set_graph(self, enabled: bool):
        self.lib._call("batch_set_graph", self.h, int(bool(enabled)))

    def graph_stats(self):
        c, r = i64(0), i64(0)
        self.lib._call("batch_graph_stats", self.h, C.byref(c), C.byref(r))
        return c.value, r.value

    def teams(self) -> np.ndarray:
        """team size of every instance in the last batched step"""
        out = np.zeros(len(self.integs), dtype=np.int64)
        self.lib._call("batch_teams", self.h, _ip(out))
        return out

    def destroy(self):
        if self.h:
            self.lib.fn["batch_destroy"](self.h)
            self.h = vp()


class Node:
    """One large branch node on the whole GPU (ttn_node_*): children of dimensions d0, d1,
    bond r, effective Hamiltonian sum_t coeff_t (op0_t (x) op1_t) A mf_t^T."""

    def __init__(self, lib: Library, op0, op1, mf, coeff=None):
        op0, op1, mf = (np.ascontiguousarray(np.asarray(x, dtype=complex)) for x in (op0, op1, mf))
        self.nt, self.d0 = op0.shape[0], op0.shape[1]
        self.d1, self.r = op1.shape[1], mf.shape[1]
        coeff = np.ones(self.nt, dtype=complex) if coeff is None else np.asarray(coeff, dtype=complex)
        self.lib = lib
        self.h = vp()
        bufs = [_cplx_in(coeff), _cplx_in(op0), _cplx_in(op1), _cplx_in(mf)]
        lib._call("node_create", lib.ctx, self.d0, self.d1, self.r, self.nt, *[_dp(b) for b in bufs],
                  C.byref(self.h))
        self.shape = (self.d0 * self.d1, self.r)

    def set(self, a):
        a = np.asarray(a, dtype=complex).reshape(self.shape)
        self.lib._call("node_set", self.h, _dp(_cplx_in(a)))

    def get(self):
        buf = np.zeros(2 * self.shape[0] * self.shape[1])
        self.lib._call("node_get", self.h, _dp(buf))
        return buf.view(np.complex128).reshape(self.shape)

    def heff(self, reps: int = 1):
        self.lib._call("node_heff", self.h, reps)
        buf = np.zeros(2 * self.shape[0] * self.shape[1])
        self.lib._call("node_get_result", self.h, _dp(buf))
        return buf.view(np.complex128).reshape(self.shape)

    def heff_timed(self, reps: int):
        self.lib._call("node_heff", self.h, reps)
        return self.last()

    def mf(self, child: int, reps: int = 1):
        dc = self.d0 if child == 0 else self.d1
        buf = np.zeros(2 * self.nt * dc * dc)
        self.lib._call("node_mf", self.h, child, reps, _dp(buf))
        return buf.view(np.complex128).reshape(self.nt, dc, dc)

    def expv(self, tau: float, krylov_dim: int = 16, tol: float = 1e-12, coeff: complex = -1j):
        na = i64(0)
        self.lib._call("node_expv", self.h, coeff.real, coeff.imag, tau, krylov_dim, tol, C.byref(na))
        return na.value

    def last(self):
        ms, fl = C.c_double(), C.c_double()
        self.lib._call("node_last", self.h, C.byref(ms), C.byref(fl))
        return ms.value, fl.value

    def destroy(self):
        if self.h:
            self.lib.fn["node_destroy"](self.h)
            self.h = vp()


SIGMA = {
    "x": np.array([[0, 1], [1, 0]], dtype=complex),
    "y": np.array([[0, -1j], [1j, 0]], dtype=complex),
    "z": np.array([[1, 0], [0, -1]], dtype=complex),
    "1": np.eye(2, dtype=complex),
}
