"""The bath discretisation of the driver front-end (ttns_lib_amd/bath.py; reference
src/spin_boson.cpp:292-365 and src/bath/bath_types/{exponential,debye,sudden}.hpp).

* against mathematics: on a weight that is a classical one the rule is the classical Gauss rule
  (Legendre: weight 1; Laguerre: e^{-x} on a window wide enough that the truncation is below
  rounding); on every bath weight the rule integrates the weight's moments (degrees up to
  2N - 1) like an independent fine quadrature; the windows satisfy their defining conditions
  (w_max = wc ln(1/tol) for s <= 1, the integrated density reaching 1 - tol for s > 1 and, at
  finite temperature, the Boltzmann-weighted density reaching 1 - tol of its integral at -w_min);
* against the reference's own printed baths (tests/golden/ref_outputs.json, from the stderr logs
  temp/errfiles_* and the identification in make_ref_outputs.py), 512 modes each:
    - zero-temperature weight J on [0, 172.69] (= 25 ln 1e3, tolerance 1e-3): the ohmic and the
      subohmic exponential baths, nodes and couplings within 1e-8 / 1e-7 relative;
    - the Debye bath on [0, 500] (its input's max frequency cutoff): 1e-7;
    - the thermofield weight (S + J) / 2 of the sudden bath at beta = 10 on [-1, 25]: 1e-7;
    - the superohmic (s = 3) bath: the largest nodes within 1e-9 (the 64 largest 1e-7), the
      smallest within 2e-2, the difference growing steadily towards w = 0 -- this
      rule integrates the weight's moments to 1e-14 and is unchanged (3e-13) on a four times
      finer discretisation, so the difference at the small nodes is the reference's own
      (modified moments against Jacobi polynomials with a 1e-10 tolerance lose accuracy where
      the x^3 weight vanishes).
  The printed exponential and Debye baths of runs whose inputs carry beta = 10 are the
  zero-temperature weight's rule; the library's discretisation follows the reference's code
  (thermofield weight whenever a temperature is given), whose thermofield branch the sudden run
  pins.  Which inputs those logs came from is not recorded beside them.
"""
import io
import json
import math
import os

import numpy as np
import pytest

from ttns_lib_amd.bath import SpectralDensity, _integrate, discretise, discretise_input, gauss_rule

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_outputs.json")


def _rule(sd, a, b, n):
    nodes, w = gauss_rule(sd.weight, a, b, n, grade_to=(0.0,) if a < 0 else ())
    o = np.argsort(np.abs(nodes), kind="stable")
    return nodes[o], np.sqrt(w[o] / math.pi)


def test_gauss_rule_reduces_to_legendre_and_laguerre():
    x, w = gauss_rule(lambda t: np.ones_like(t), -1.0, 1.0, 24)
    xl, wl = np.polynomial.legendre.leggauss(24)
    assert np.abs(x - xl).max() < 1e-13 and np.abs(w - wl).max() < 1e-13
    x, w = gauss_rule(lambda t: np.exp(-t), 0.0, 200.0, 20)
    xl, wl = np.polynomial.laguerre.laggauss(20)
    assert np.abs(x / xl - 1).max() < 1e-11 and np.abs(w / wl - 1).max() < 1e-9


@pytest.mark.parametrize("kind,s,beta", [("exponential", 0.5, math.inf), ("exponential", 1.0, 10.0),
                                         ("exponential", 3.0, math.inf), ("debye", 1.0, 10.0),
                                         ("sudden", 1.0, 10.0)])
def test_rule_integrates_the_weights_moments(kind, s, beta):
    sd = SpectralDensity(kind, 2.0, 25.0, s, beta)
    a, b = sd.bounds(1e-3)
    if kind == "debye":
        b = 500.0
    n = 64
    x, w = gauss_rule(sd.weight, a, b, n, grade_to=(0.0,) if a < 0 else ())
    c, h = 0.5 * (a + b), 0.5 * (b - a)
    for k in (0, 1, 7, 31, 2 * n - 1):
        # Legendre polynomials of the scaled variable keep the moments well conditioned
        P = np.polynomial.legendre.Legendre.basis(k)
        want = _integrate(lambda t: sd.weight(t) * P((t - c) / h), a, b, npanel=3000)
        got = np.sum(w * P((x - c) / h))
        scale = _integrate(lambda t: np.abs(sd.weight(t)), a, b, npanel=3000)
        assert abs(got - want) < 1e-11 * scale, (k, got, want)


def test_windows_meet_their_definitions():
    tol = 1e-3
    sd = SpectralDensity("exponential", 2.0, 25.0, 1.0)
    assert sd.bounds(tol) == (0.0, 25.0 * math.log(1e3))
    sd3 = SpectralDensity("exponential", 2.0, 25.0, 3.0)
    _, b = sd3.bounds(tol)
    total = math.gamma(3.0) * 2.0 * 25.0 / 2.0 * math.pi
    assert abs(_integrate(sd3.density, 0.0, b) / ((1 - tol) * total) - 1) < 1e-10
    sdt = SpectralDensity("sudden", 2.0, 25.0, 1.0, beta=10.0)
    a, b = sdt.bounds(tol)
    f = lambda x: sdt.density(x) * np.exp(-10.0 * x)  # noqa: E731
    assert b == 25.0 and a < 0
    assert abs(_integrate(f, 0.0, -a) / ((1 - tol) * _integrate(f, 0.0, b)) - 1) < 1e-10


def _printed(name):
    with open(GOLDEN) as f:
        run = json.load(f)[name]
    return np.array(run["wk"]), np.array(run["gk"])


@pytest.mark.parametrize("name,kind,s,beta,window,bw,bg", [
    ("ohmic_exponential", "exponential", 1.0, math.inf, (0.0, 25.0 * math.log(1e3)), 1e-8, 1e-8),
    ("subohmic_exponential_t0", "exponential", 0.5, math.inf, (0.0, 25.0 * math.log(1e3)), 1e-7, 1e-7),
    ("ohmic_debye", "debye", 1.0, math.inf, (0.0, 500.0), 1e-7, 1e-7),
    ("ohmic_sudden", "sudden", 1.0, 10.0, (-1.0, 25.0), 1e-9, 1e-7)])
def test_rule_reproduces_the_references_printed_baths(name, kind, s, beta, window, bw, bg):
    wr, gr = _printed(name)
    w, g = _rule(SpectralDensity(kind, 2.0, 25.0, s, beta), *window, len(wr))
    assert np.abs(w / wr - 1).max() < bw
    assert np.abs(g / gr - 1).max() < bg


def test_superohmic_bath_within_the_references_accuracy():
    wr, gr = _printed("superohmic_exponential")
    sd = SpectralDensity("exponential", 2.0, 25.0, 3.0)
    w, g = _rule(sd, 0.0, sd.bounds(1e-3)[1], len(wr))
    rel = np.abs(w / wr - 1)
    assert rel[-8:].max() < 1e-9 and rel[-64:].max() < 1e-7 and rel[256:].max() < 1e-6 and rel.max() < 2e-2
    assert np.all(np.diff(rel[:100]) < 0)   # the difference shrinks away from w = 0
    # refining the discretisation four times moves this rule by rounding only
    n1, _ = gauss_rule(sd.weight, 0.0, sd.bounds(1e-3)[1], 512, panels=2048)
    assert np.abs(np.sort(n1) / np.sort(w) - 1).max() < 1e-11


def test_driver_discretises_the_inputs_bath():
    cfg = {"nmodes": 48, "bath": {"type": "sudden", "wc": 25.0, "alpha": 2.0, "s": 1, "beta": 10},
           "bathintegrationerrortolerance": 1e-3, "negativefrequencycutoff": -1.0}
    w, g = discretise_input(cfg)
    w2, g2 = _rule(SpectralDensity("sudden", 2.0, 25.0, 1.0, 10.0), -1.0, 25.0, 48)
    assert np.array_equal(w, w2) and np.array_equal(g, g2)
    assert np.all(np.diff(np.abs(w)) >= 0)
    # zero temperature ignores the negative cutoff; a max cutoff replaces the window
    cfg0 = {"nmodes": 32, "bath": {"type": "exponential", "wc": 5.0, "alpha": 0.1, "s": 1},
            "maxfrequencycutoff": 30.0, "negativefrequencycutoff": -2.0}
    w0, _ = discretise_input(cfg0)
    assert w0.min() > 0 and w0.max() < 30.0
    w1, g1 = discretise(SpectralDensity("exponential", 0.1, 5.0, 1.0), 32, tol=1e-5)
    assert w1.max() < 5.0 * math.log(1e5)
