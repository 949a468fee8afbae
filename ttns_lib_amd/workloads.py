"""Spin-boson workloads: spectral densities, bath discretisation and the benchmark sets.

Spectral densities follow the reference's bath types:
  exponential  J(w) = (pi alpha / 2) wc (|w|/wc)^s exp(-|w|/wc) sign(w)   (src/bath/bath_types/exponential.hpp:36-39)
  debye        J(w) = alpha wc^2 w / (w^2 + wc^2)                          (src/bath/bath_types/debye.hpp:35-38)
At finite temperature the bath is folded into a zero-temperature one over positive and
negative frequencies with the density (S + J)/2, S(w) = J(w) coth(beta w / 2)
(src/spin_boson.cpp:311-328).  The discretisation here is Gauss-Legendre quadrature of that
density on [0, wmax] (T = 0) or [-wmax, wmax] (T > 0), g_k = sqrt(weight_k J_eff(w_k) / pi)
(spin_boson.cpp:362), modes sorted by |w| (spin_boson.cpp:353).  The reference builds its
nodes from orthogonal polynomials of the density itself; the arrays produced here are only
ever inputs, fed identically to the GPU path and to the oracle.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

from .capi import SpinBoson

GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden")


def j_exponential(w, alpha, wc, s=1.0):
    w = np.asarray(w, dtype=float)
    return np.pi * alpha / 2.0 * wc * (np.abs(w) / wc) ** s * np.exp(-np.abs(w) / wc) * np.sign(w)


def j_debye(w, alpha, wc):
    w = np.asarray(w, dtype=float)
    return alpha * wc * wc * w / (w * w + wc * wc)


def effective_density(J, beta):
    """(S + J)/2 with S = J coth(beta w/2); beta = inf gives J on w > 0."""
    def f(w):
        w = np.asarray(w, dtype=float)
        if np.isinf(beta):
            return np.where(w > 0, J(w), 0.0)
        x = beta * w / 2.0
        with np.errstate(divide="ignore", invalid="ignore"):
            coth = np.where(np.abs(x) > 1e-8, 1.0 / np.tanh(x), 0.0)
        S = np.where(np.abs(x) > 1e-8, J(w) * coth, 0.0)
        return 0.5 * (S + J(w))
    return f


def discretise(J, nmodes, wmax, beta=np.inf):
    lo = 0.0 if np.isinf(beta) else -wmax
    x, wt = np.polynomial.legendre.leggauss(nmodes)
    w = lo + (x + 1.0) * (wmax - lo) / 2.0
    wt = wt * (wmax - lo) / 2.0
    dens = effective_density(J, beta)(w)
    g = np.sqrt(np.maximum(dens, 0.0) * wt / np.pi)
    order = np.argsort(np.abs(w), kind="stable")
    return w[order], g[order]


def ohmic_bath(nmodes, alpha, wc, beta=np.inf, wmax_factor=8.0):
    J = lambda w: j_exponential(w, alpha, wc, 1.0)  # noqa: E731
    return discretise(J, nmodes, wmax_factor * wc, beta)


def debye_bath(nmodes, alpha, wc, beta=np.inf, wmax_factor=20.0):
    J = lambda w: j_debye(w, alpha, wc)  # noqa: E731
    return discretise(J, nmodes, wmax_factor * wc, beta)


@dataclass
class Instance:
    model: SpinBoson
    label: str
    dt: float
    krylov_dim: int = 12


def c0_instance() -> Instance:
    """Plain small: Ohmic exponential alpha 0.1, wc 5, T = 0, 16 modes, boson dim 10,
    bond dim 8 (binary tree of depth 5), dt 0.05, Krylov dim 12."""
    w, g = ohmic_bath(16, 0.1, 5.0)
    return Instance(SpinBoson(wk=w, gk=g, boson_dim=[10] * 16, nspf=8, nspf_lower=8), "c0", 0.05, 12)


def c1_instance() -> Instance:
    """Single larger run: Debye alpha 0.5, wc 2, beta 10 (thermofield: 128 modes over
    positive and negative frequencies), boson dim 16, bond dim 32, dt 0.025."""
    w, g = debye_bath(128, 0.5, 2.0, beta=10.0)
    return Instance(SpinBoson(wk=w, gk=g, boson_dim=[16] * 128, nspf=32, nspf_lower=32), "c1", 0.025, 12)


# the expansion knobs of the single larger run: spawning 1e-8, unoccupied threshold 1e-4, at least
# two unoccupied functions (the reference driver's keys "spawning parameter", "unoccupied
# threshold", "minimum unoccupied", src/spin_boson.cpp:258-278)
C1_EXPANSION = dict(spawning_threshold=1e-8, unoccupied_threshold=1e-4, minimum_unoccupied=2)


def c1_expansion_instance() -> Instance:
    """The single larger run as configured, subspace expansion on: the tree of c1_instance with
    initial bond dimension 8 and capacity 32 (leaves capped at their 16 states), the bonds grown by
    the expansion integrator with the C1_EXPANSION knobs."""
    w, g = debye_bath(128, 0.5, 2.0, beta=10.0)
    return Instance(SpinBoson(wk=w, gk=g, boson_dim=[16] * 128, nspf=8, nspf_lower=8, nspf_cap=32, nspf_lower_cap=32),
                    "c1x", 0.025, 12)


def c4_instance() -> Instance:
    """Adjoint config: Ohmic alpha 0.2, wc 5, T = 0, 64 modes, boson dim 12, bond dim 24."""
    w, g = ohmic_bath(64, 0.2, 5.0)
    return Instance(SpinBoson(wk=w, gk=g, boson_dim=[12] * 64, nspf=24, nspf_lower=24), "c4", 0.05, 12)


C3_FILE = os.path.join(GOLDEN, "c3_instances.json")


def make_c3_spec(seed: int = 2024, n: int = 64):
    """64 uneven instances: alpha ~ U[0.05, 1], wc in {1, 2, 5}, beta in {1, 10, inf},
    32-128 modes, boson dims 8-20, bond dims 16-48."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        alpha = float(rng.uniform(0.05, 1.0))
        wc = float(rng.choice([1.0, 2.0, 5.0]))
        beta = [1.0, 10.0, float("inf")][int(rng.integers(0, 3))]
        nmodes = int(rng.integers(32, 129))
        d = int(rng.integers(8, 21))
        r = int(rng.integers(16, 49))
        w, g = ohmic_bath(nmodes, alpha, wc, beta)
        out.append(dict(index=i, alpha=alpha, wc=wc, beta=beta if np.isfinite(beta) else "inf", nmodes=nmodes,
                        boson_dim=d, bond_dim=r, wk=[float(x) for x in w], gk=[float(x) for x in g]))
    return out


def c3_instances(path: str = C3_FILE):
    with open(path) as f:
        spec = json.load(f)
    insts = []
    for s in spec:
        m = SpinBoson(wk=s["wk"], gk=s["gk"], boson_dim=[s["boson_dim"]] * s["nmodes"], nspf=s["bond_dim"],
                      nspf_lower=s["bond_dim"], seed=s["index"])
        insts.append(Instance(m, f"c3[{s['index']}]", 0.05, 12))
    return insts


REF_FILE = os.path.join(GOLDEN, "ref_outputs.json")


def reference_runs(lib, count: int = 64, path: str = REF_FILE):
    """The reference's own 512-mode example runs (exact discretised baths and input keys of
    example_outputs/{beta_10/{superohmic,ohmic},zero_temperature/subohmic}), built as its driver
    builds them (partition_modes, mode blocks; src/spin_boson.cpp:418-538) at their initial bond
    dimensions 4-8, dt 0.005: `count` instances cycling over the five runs, each copy with its own
    seed for the random unoccupied functions."""
    from .capi import driver_blocks

    with open(path) as f:
        runs = json.load(f)
    names = sorted(runs)
    out = []
    for i in range(count):
        name = names[i % len(names)]
        run, cfg = runs[name], runs[name]["input"]
        sizes, nmax = driver_blocks(lib, run["wk"], ecut=3000.0, high_frequency_bound=cfg["high frequency bound"],
                                    target_dimension=cfg["target dimension"],
                                    maximum_dimension=cfg["maximum dimension"],
                                    maximum_dimension_high_frequency=cfg["maximum dimension high frequency"])
        m = SpinBoson(wk=np.array(run["wk"]), gk=np.array(run["gk"]), block_sizes=sizes, block_nmax=nmax,
                      nspf=cfg["nspf"], nspf_lower=cfg["nspf lower"], seed=cfg["seed"] + i // len(names),
                      hsys=np.array(cfg["Hsys"], dtype=complex), scoup=np.array(cfg["Scoup"], dtype=complex),
                      psi0=np.array(cfg["psi0"], dtype=complex))
        out.append(Instance(m, f"{name}#{i // len(names)}", cfg["dt"], cfg["krylov dimension"]))
    return out
