"""Bath discretisation of the spin-boson driver: the continuous spectral densities of the
reference's input files turned into the N harmonic modes (w_k, g_k) of the star Hamiltonian
(reference src/spin_boson.cpp:292-365, bath types src/bath/bath_types/{exponential,debye,
sudden}.hpp).

The reference builds the polynomials orthonormal with respect to the bath's weight on its
frequency window and takes their Gauss rule: w_k are the nodes, g_k = sqrt(W_k / pi) with W_k
the weights, ordered by |w_k|.  The weight is J(w) on [0, w_max] at zero temperature and, at
inverse temperature beta, the thermofield weight (S(w) + J(w)) / 2 = J(w) / (1 - exp(-beta w))
on [w_min, w_max] with w_min < 0.  The window follows the reference's frequency_bounds: w_max from
the integrated density of frequencies reaching 1 - tol of its total (tol = 1e-5, the
"bath integration error tolerance"), or its closed form, and w_min where the Boltzmann-weighted
density reaches 1 - tol of its integral over [0, w_max].

The Gauss rule here is computed by the discretisation method (Gautschi, "Orthogonal Polynomials:
Computation and Approximation", 2004, sec. 2.2): the weight is discretised by composite
Gauss-Legendre panels graded towards zero and the window's ends, the recurrence coefficients of
the discrete measure come from the Lanczos process with full reorthogonalisation, and the nodes
and weights from the symmetric tridiagonal (Jacobi) matrix (Golub and Welsch 1969).  The
reference obtains the same rule through modified moments against shifted Jacobi polynomials (the
un-vendored orthopol library, no pinned version); the rule itself is unique, so both agree to the
accuracy of their computations.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import eigh_tridiagonal
from scipy.special import gamma


class SpectralDensity:
    """J(w) and the density of frequencies of one bath type; beta = inf for zero temperature."""

    def __init__(self, kind: str, alpha: float, wc: float, s: float = 1.0, beta: float = math.inf):
        if kind not in ("exponential", "debye", "sudden"):
            raise ValueError(f"bath type {kind!r} is not supported (exponential, debye, sudden)")
        self.kind, self.alpha, self.wc, self.s, self.beta = kind, float(alpha), float(wc), float(s), float(beta)

    @classmethod
    def from_input(cls, bath: dict):
        """the "bath" object of a reference input file (keys type, alpha, wc, s, beta or T)"""
        b = {k.replace(" ", "").lower(): v for k, v in bath.items()}
        beta = math.inf
        if "beta" in b:
            beta = float(b["beta"])
        elif "t" in b and float(b["t"]) > 1e-14:
            beta = 1.0 / float(b["t"])
        return cls(b["type"], b["alpha"], b["wc"], b.get("s", 1.0), beta)

    @property
    def finite_temperature(self) -> bool:
        return math.isfinite(self.beta)

    def J(self, w):
        w = np.asarray(w, dtype=float)
        a, wc, s = self.alpha, self.wc, self.s
        x = np.abs(w)
        if self.kind == "exponential":
            v = math.pi * a / 2.0 * wc * (x / wc) ** s * np.exp(-x / wc)
        elif self.kind == "debye":
            return a * wc * wc * w / (w * w + wc * wc)
        else:
            v = np.where(x < wc, math.pi * a * gamma(s + 1) / 2.0 * wc * (x / wc) ** s, 0.0)
        return np.where(w < 0, -v, v)

    def density(self, w):
        """density of frequencies (exponential.hpp:93-97, debye.hpp:66, sudden.hpp:66-70)"""
        w = np.asarray(w, dtype=float)
        a, wc, s = self.alpha, self.wc, self.s
        x = np.abs(w)
        if self.kind == "debye":
            return a * wc * wc / (2.0 * (w * w + wc * wc))
        if s < 1:
            return self.J(w)
        if self.kind == "exponential":
            return math.pi / 2.0 * a * (x / wc) ** (s - 1.0) * np.exp(-x / wc)
        v = np.where(x < wc, math.pi * a * gamma(s + 1) / 2.0 * (x / wc) ** (s - 1), 0.0)
        return np.where(w < 0, -v, v)

    def weight(self, w):
        """the Gauss weight: J at zero temperature, J / (1 - exp(-beta w)) otherwise"""
        w = np.asarray(w, dtype=float)
        if not self.finite_temperature:
            return self.J(w)
        bw = self.beta * w
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            # 1 / (1 - exp(-x)) = 1 / x + 1 / 2 + O(x) for small x (the quadrature never samples 0)
            return np.where(np.abs(bw) > 1e-8, self.J(w) / -np.expm1(-bw), self.J(w) / bw + 0.5 * self.J(w))

    def bounds(self, tol: float = 1e-5):
        """[w_min, w_max] (frequency_bounds with the thermofield window at finite temperature)"""
        a, wc, s = self.alpha, self.wc, self.s
        if self.kind == "exponential":
            if s > 1:
                total = gamma(s) * a * wc / 2.0 * math.pi
                wmax = _solve_cumulative(self.density, (1.0 - tol) * total, wc * math.log(1.0 / tol))
            else:
                wmax = wc * math.log(1.0 / tol)
        elif self.kind == "debye":
            wmax = wc * math.tan((1.0 - tol) * math.pi / 2.0)
        else:
            wmax = wc
        wmin = 0.0
        if self.finite_temperature:
            f = lambda x: self.density(x) * np.exp(-self.beta * np.asarray(x))  # noqa: E731
            total = _integrate(f, 0.0, wmax)
            wmin = -_solve_cumulative(f, (1.0 - tol) * total, wmax / 2.0)
        return wmin, wmax


def _gl_panels(a: float, b: float, edges, order: int = 24):
    """composite Gauss-Legendre nodes and weights on [a, b] with panel edges `edges` (sorted)"""
    x0, w0 = np.polynomial.legendre.leggauss(order)
    xs, ws = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        if hi <= lo:
            continue
        xs.append(0.5 * (hi - lo) * x0 + 0.5 * (hi + lo))
        ws.append(0.5 * (hi - lo) * w0)
    return np.concatenate(xs), np.concatenate(ws)


def _graded_edges(a: float, b: float, npanel: int, grade_to=(), depth: int = 40):
    """uniform panel edges on [a, b] plus geometric refinement towards each point of grade_to"""
    e = list(np.linspace(a, b, npanel + 1))
    h = (b - a) / npanel
    for p in grade_to:
        for k in range(1, depth):
            d = h * 0.5 ** k
            for q in (p - d, p + d):
                if a < q < b:
                    e.append(q)
    return np.unique(np.array(e))


def _integrate(f, a: float, b: float, npanel: int = 400) -> float:
    x, w = _gl_panels(a, b, _graded_edges(a, b, npanel, grade_to=(a, b)))
    return float(np.sum(w * f(x)))


def _solve_cumulative(f, target: float, x0: float, xtol: float = 1e-12) -> float:
    """x > 0 with integral_0^x f = target (f >= 0), by bracketing and bisection-safeguarded
    Newton steps on the cumulative integral"""
    lo, hi = 0.0, max(x0, 1e-12)
    while _integrate(f, 0.0, hi) < target:
        lo, hi = hi, 2.0 * hi
    x = 0.5 * (lo + hi)
    for _ in range(200):
        F = _integrate(f, 0.0, x) - target
        if F > 0:
            hi = x
        else:
            lo = x
        d = float(f(np.array([x]))[0])
        xn = x - F / d if d > 0 else 0.5 * (lo + hi)
        if not (lo < xn < hi):
            xn = 0.5 * (lo + hi)
        if abs(xn - x) <= xtol * max(1.0, abs(x)):
            return xn
        x = xn
    return x


def gauss_rule(weight, a: float, b: float, n: int, panels: int = 0, grade_to=()):
    """n-point Gauss rule of the weight on [a, b] by the discretisation method: nodes, weights"""
    panels = panels or max(64, n)
    x, w = _gl_panels(a, b, _graded_edges(a, b, panels, grade_to=(a, b) + tuple(grade_to)))
    rho = np.maximum(weight(x), 0.0) * w
    keep = rho > 0
    x, rho = x[keep], rho[keep]
    mu0 = rho.sum()
    # Lanczos on diag(x) from the start vector sqrt(rho / mu0), fully reorthogonalised
    m = x.size
    if n > m:
        raise ValueError("the discretisation has fewer points than the rule")
    Q = np.zeros((n, m))
    al, be = np.zeros(n), np.zeros(max(n - 1, 0))
    q = np.sqrt(rho / mu0)
    for k in range(n):
        Q[k] = q
        v = x * q
        al[k] = q @ v
        v = v - al[k] * q - (be[k - 1] * Q[k - 1] if k > 0 else 0.0)
        for _ in range(2):
            v -= Q[:k + 1].T @ (Q[:k + 1] @ v)
        if k + 1 < n:
            be[k] = np.linalg.norm(v)
            q = v / be[k]
    nodes, vecs = eigh_tridiagonal(al, be)
    return nodes, mu0 * vecs[0] ** 2


def discretise(density: SpectralDensity, n: int, tol: float = 1e-5, panels: int = 0):
    """(w_k, g_k), k = 1..n, ordered by |w_k| (src/spin_boson.cpp:292-365)"""
    wmin, wmax = density.bounds(tol)
    nodes, weights = gauss_rule(density.weight, wmin, wmax, n, panels, grade_to=(0.0,) if wmin < 0 else ())
    order = np.argsort(np.abs(nodes), kind="stable")
    return nodes[order], np.sqrt(weights[order] / math.pi)


def discretise_input(cfg: dict):
    """(w_k, g_k) of a reference input file (keys without spaces, lower case; driver.load_input):
    nmodes, bath, and the optional bath integration error tolerance (default 1e-5), max
    frequency cutoff and negative frequency cutoff (src/spin_boson.cpp:192-222, 297-331)."""
    sd = SpectralDensity.from_input(cfg["bath"])
    n = int(cfg["nmodes"])
    tol = float(cfg.get("bathintegrationerrortolerance", 1e-5))
    if "maxfrequencycutoff" in cfg:
        wmax = float(cfg["maxfrequencycutoff"])
        wmin = -wmax
    else:
        wmin, wmax = sd.bounds(tol)
    if sd.finite_temperature:
        wmin = float(cfg.get("negativefrequencycutoff", wmin))
    else:
        wmin = 0.0
    nodes, weights = gauss_rule(sd.weight, wmin, wmax, n, grade_to=(0.0,) if wmin < 0 else ())
    order = np.argsort(np.abs(nodes), kind="stable")
    return nodes[order], np.sqrt(weights[order] / math.pi)

