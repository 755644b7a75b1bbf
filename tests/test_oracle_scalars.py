"""The oracle's scalar physics against the reference core's point values
(tests/golden/scalars.npz, written by tests/golden/make_golden.py from the
reference C core) and against closed forms that a plausible mistake breaks."""
import math

import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden


@pytest.fixture(scope="module")
def g():
    return load_golden("scalars")


def test_bernoulli_matches_reference(g):
    np.testing.assert_allclose(O.bernoulli(g["B_x"]), g["B"], rtol=1e-14, atol=0)


def test_bernoulli_closed_forms():
    x = np.logspace(-5, 2.8, 200)
    B = O.bernoulli(x)
    # B(-x) = B(x) + x (evaluated for x > 0, where the sum does not cancel)
    np.testing.assert_allclose(O.bernoulli(-x), B + x, rtol=1e-12)
    # B(x) e^x = B(-x)
    xs = x[x < 50]
    np.testing.assert_allclose(O.bernoulli(xs) * np.exp(xs), O.bernoulli(-xs), rtol=1e-12)
    assert O.bernoulli(np.array([0.0]))[0] == 1.0
    np.testing.assert_allclose(O.bernoulli(np.array([1.0]))[0], 1.0 / (math.e - 1.0), rtol=1e-15)
    np.testing.assert_allclose(O.bernoulli(np.array([-1.0]))[0], math.e / (math.e - 1.0), rtol=1e-15)
    assert O.bernoulli(np.array([800.0]))[0] == 0.0 and O.bernoulli(np.array([-800.0]))[0] == 800.0


def test_silicon_parameters_match_reference(g):
    m = O.silicon(300.0)
    VT, Eg, chi, eps_r, Nc, Nv, ni = g["mat0"][:7]
    assert m["VT"] == pytest.approx(VT, rel=1e-15)
    assert m["Eg"] == pytest.approx(Eg, rel=1e-15)
    assert m["Nc"] == pytest.approx(Nc, rel=1e-13) and m["Nv"] == pytest.approx(Nv, rel=1e-13)
    assert m["ni"] == pytest.approx(ni, rel=1e-12)
    # n_i of Si at 300 K: about 1e10 cm^-3 with these band parameters (validate.py section 18)
    assert 0.5e16 < m["ni"] < 2e16


def test_masetti_mobility_matches_reference(g):
    m = O.silicon(300.0)
    np.testing.assert_allclose(O.mob_low(m, g["N"], 0), g["mun"], rtol=1e-13)
    np.testing.assert_allclose(O.mob_low(m, g["N"], 1), g["mup"], rtol=1e-13)
    # limits: lattice mobility at low doping, mu_min1 + mu_1 ordering at high doping
    assert O.mob_low(m, 1.0, 0) == pytest.approx(0.1417 - 0.00434 * 0 + 0.0, rel=2e-2)
    assert O.mob_low(m, 1e27, 0) < 0.01


def test_bgn_matches_reference(g):
    m = O.silicon(300.0)
    np.testing.assert_allclose(O.bgn_half(m, g["N"]), g["bgn"], rtol=1e-13, atol=1e-300)
    # Klaassen form at N = N_ref: dEg = E (0 + sqrt(C)) -> half gap / VT
    assert O.bgn_half(m, np.array([1.3e23]))[0] == pytest.approx(0.5 * 6.92e-3 * math.sqrt(0.5) / m["VT"], rel=1e-14)


def test_recombination_matches_reference(g):
    m = O.silicon(300.0)
    m["tau"] = (float(g["mat0"][13]), float(g["mat0"][14]))    # the fixture device sets tau 1e-7 s
    r = g["recomb"]
    R, dRn, dRp = O.recomb(m, 1e16, 0.5, r[:, 0], r[:, 1])
    np.testing.assert_allclose(R, r[:, 2], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(dRn, r[:, 3], rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(dRp, r[:, 4], rtol=1e-12, atol=1e-300)


def test_recombination_closed_forms():
    m = O.silicon(300.0)
    ni = m["ni"]
    # equilibrium np = ni^2: no net recombination
    n = np.logspace(5, 25, 9)
    R, _, _ = O.recomb(m, ni, 1.0, n, ni * ni / n)
    assert np.all(np.abs(R) <= 1e-12 * (n + ni * ni / n) / 1e-5)
    # low injection in p-type: SRH -> (n - n0)/tau_n (plus Auger/radiative terms)
    p = 1e23
    dn = 1e10
    R, _, _ = O.recomb(m, ni, 1.0, ni * ni / p + dn, p)
    tau = m["tau"][0]
    expect = dn * p / (tau * (ni * ni / p + dn + ni) + tau * (p + ni)) + (m["Cn"] * (ni * ni / p + dn) + m["Cp"] * p + m["Brad"]) * dn * p
    assert float(R) == pytest.approx(expect, rel=1e-6)
    # derivative check by finite differences
    n0, p0 = 3e18, 2e17
    R0, dRn, dRp = O.recomb(m, ni, 0.7, n0, p0)
    h = 1e-6
    assert (O.recomb(m, ni, 0.7, n0 * (1 + h), p0)[0] - R0) / (n0 * h) == pytest.approx(dRn, rel=1e-5)
    assert (O.recomb(m, ni, 0.7, n0, p0 * (1 + h))[0] - R0) / (p0 * h) == pytest.approx(dRp, rel=1e-5)


def test_velocity_saturation_closed_forms():
    mu0, vs, beta = 0.1, 1e5, 2.0
    assert O.mob_hf(mu0, 0.0, vs, beta) == mu0
    E = np.logspace(3, 8, 20)
    mu = O.mob_hf(mu0, E, vs, beta)
    np.testing.assert_allclose(mu, mu0 / np.sqrt(1 + (mu0 * E / vs) ** 2), rtol=1e-14)
    assert mu[-1] * E[-1] == pytest.approx(vs, rel=1e-3)      # v -> v_sat
    assert O.heat_field(0.1, -0.2, 1e-9) == 0.0
    assert O.heat_field(0.1, 0.1, 1e-9) == pytest.approx(0.1 / 1e-9)   # drift: F = |E|
