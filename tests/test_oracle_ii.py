"""The oracle's impact-ionisation generation (ii_update) against the
reference core's, dumped at perturbed reverse-bias states of a pn diode and
a MOSFET (tests/golden/k_ii_*.npz), plus closed forms of the ionisation
coefficient and of the node field."""
import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device

FIXTURES = ["k_ii_diode_6x5x16", "k_ii_mosfet_20x14x9"]


@pytest.fixture(scope="module", params=FIXTURES)
def case(request):
    g = load_golden(request.param)
    return g, oracle_device(g)


def test_fixture_has_avalanche(case):
    g, _ = case
    assert (g["gii"] > 0).sum() > 50 and g["gii"].max() > 0


def test_ii_update_matches_reference(case):
    g, dev = case
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    np.testing.assert_allclose(n, g["n"], rtol=1e-12)
    G = O.ii_update(dev, g["psi"], n, p, g["fmu"])
    ref = g["gii"]
    # the reference is built with -ffast-math: compare to 1e-9 of each value
    # plus 1e-12 of the largest
    assert np.all(np.abs(G - ref) <= 1e-9 * np.abs(ref) + 1e-12 * ref.max()), np.max(np.abs(G - ref))
    np.testing.assert_array_equal(G > 0, ref > 0)


def test_alpha_closed_forms():
    m = O.silicon()
    assert m["ii_g"] == pytest.approx(1.0, abs=1e-15)          # hw correction is 1 at 300 K
    F = np.array([0.0, 5e4, 1e5, 1e5 * (1 + 1e-12), 2e7, 4e7, 1e8])
    for c in (0, 1):
        a = O.ii_alpha(m, F, c)
        assert a[0] == 0 and a[1] == 0 and a[2] == 0            # at or below 1 kV/cm: none
        assert a[3] == 0                                         # b/F > 700: underflow cut to 0
        assert np.all(np.diff(a[3:]) > 0)                        # increasing with field
    # electrons: one coefficient pair on both sides of E0 (continuous there)
    lo, hi = O.ii_alpha(m, np.array([4e7 * (1 - 1e-12), 4e7]), 0)
    assert lo == pytest.approx(hi, rel=1e-9)
    # holes: at E0 the high-field pair takes over: alpha(E0) = 6.71e7 exp(-1.693e8/4e7)
    assert O.ii_alpha(m, np.array([4e7]), 1)[0] == pytest.approx(6.71e7 * np.exp(-1.693e8 / 4e7), rel=1e-14)
    # saturation: alpha -> a for F >> b
    assert O.ii_alpha(m, np.array([1e14]), 0)[0] == pytest.approx(7.03e7, rel=1e-5)


def test_node_field_of_linear_potential():
    """A potential linear in z has E = -slope at every semiconductor node
    (central differences are exact; one-sided at the contact faces)."""
    g = load_golden("k_ii_diode_6x5x16")
    dev = oracle_device(g)
    _, _, l = dev.ijl()
    slope = 2.5e6
    psi = slope * dev.zs[l]
    n, p = dev.carriers(psi, psi, psi)
    Jn, Jp, E = O.node_field_current(dev, psi, n, p, g["fmu"])
    np.testing.assert_allclose(E[2], -slope, rtol=1e-12)
    np.testing.assert_allclose(E[:2], 0.0, atol=1e-9)
