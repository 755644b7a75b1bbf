"""The oracle's Lombardi surface mobility (models bit 16) and interface
distances against the reference core's at perturbed solved states of two
MOSFETs (tests/golden/k_lb_*.npz), plus closed forms of the Lombardi terms."""
import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device

FIXTURES = ["k_lb_mosfet_12x8x6", "k_lb_mosfet_20x14x9"]


@pytest.fixture(scope="module", params=FIXTURES)
def case(request):
    g = load_golden(request.param)
    return g, oracle_device(g)


def test_wall_distance_matches_reference(case):
    g, dev = case
    qd = O.wall_distance(dev)
    np.testing.assert_allclose(qd, g["qd"], rtol=1e-15, atol=0)
    assert (g["qd"] < 1e-8).sum() > 10            # the gate interface is resolved


def test_face_mobility_with_lombardi_matches_reference(case):
    g, dev = case
    assert dev.models & 16
    fmu = O.face_mobility(dev, g["psi"], g["phin"], g["phip"])
    for q in range(6):
        np.testing.assert_allclose(fmu[q], g["fmu"][q], rtol=1e-12, atol=0)
    # the surface term lowers the mobility at the interface, and only there
    plain = dev.models
    dev.models = plain & ~16
    bulk = O.face_mobility(dev, g["psi"], g["phin"], g["phip"])
    dev.models = plain
    near = O.wall_distance(dev) < 1e-8
    sel = np.zeros(dev.N, bool)
    sel[:] = near
    assert np.all(fmu[0][sel] <= bulk[0][sel] * (1 + 1e-15))
    assert np.any(fmu[0][sel] < 0.99 * bulk[0][sel])


def test_lombardi_closed_forms():
    m = O.silicon()
    mub = 0.14
    # far from an interface (Dw -> 0) the bulk mobility is unchanged
    assert O.lombardi(m, mub, 1e7, 1e23, 0.0, 0) == pytest.approx(mub, rel=1e-15)
    # at an interface, Matthiessen's rule with the acoustic-phonon and
    # surface-roughness terms: 1/mu = 1/mub + 1e4 (1/mu_ac + 1/mu_sr), cm units
    F, N = 1e5, 1e17                                   # V/cm, cm^-3
    mu_ac = 4.75e7 / F + 580.0 * N ** 0.125 / np.cbrt(F)
    mu_sr = 5.82e14 / F ** 2
    want = 1.0 / (1.0 / mub + 1e4 * (1.0 / mu_ac + 1.0 / mu_sr))
    assert O.lombardi(m, mub, F * 1e2, N * 1e6, 1.0, 0) == pytest.approx(want, rel=1e-14)
    # the field and doping are floored at 1 V/cm and 1 cm^-3
    assert O.lombardi(m, mub, 0.0, 0.0, 1.0, 1) == pytest.approx(O.lombardi(m, mub, 1e2, 1e6, 1.0, 1), rel=1e-15)
    # a stronger normal field lowers the surface mobility
    assert O.lombardi(m, mub, 1e8, 1e23, 1.0, 0) < O.lombardi(m, mub, 1e7, 1e23, 1.0, 0)


def test_entry_point_values_match_reference():
    """api3_lombardi of the reference for silicon at 300 and 400 K over field,
    doping, distance factor and carrier (tests/golden/surface_models.npz)."""
    g = load_golden("surface_models")
    for r, T in enumerate((300.0, 400.0)):
        m = O.silicon(T)
        got = [O.lombardi(m, mub, F, Nt, Dw, int(h)) for mub, F, Nt, Dw, h in g["lb_in"]]
        np.testing.assert_allclose(got, g["lombardi"][r], rtol=1e-13, atol=0)
