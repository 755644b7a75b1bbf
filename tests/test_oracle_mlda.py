"""The oracle's MLDA quantum-confinement terms (models bit 3) against the
reference core's (tests/golden/k_mlda_*.npz; the fixtures' assembled
operators are checked by test_oracle_kernels), plus closed forms of the
MLDA population factor."""
import math

import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device


@pytest.mark.parametrize("name", ["k_mlda_mosfet_12x8x6", "k_mlda_mosfet_20x14x9"])
def test_confinement_terms_match_reference(name):
    g = load_golden(name)
    dev = oracle_device(g)
    assert dev.models & 8
    np.testing.assert_allclose(dev.qcn, g["qcn"], rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(dev.qcp, g["qcp"], rtol=1e-12, atol=1e-15)
    assert (g["qcn"] < -0.1).sum() > 5                 # the gate interface confines carriers
    # the confinement lowers the effective densities of states only
    assert np.all(dev.qcn <= 0) and np.all(dev.qcp <= 0)


def test_mlda_closed_forms():
    lam = 2e-9
    # far from the wall the density is the bulk one; at the wall it vanishes
    assert O.mlda_pop(1e-6, 1e-6 + 1e-9, lam) == pytest.approx(1.0, abs=1e-12)
    assert O.mlda_pop(0.0, 0.0, lam) == 0.0
    # a cell average equals the exact integral of 1 - exp(-(z/lam)^2) / (b - a)
    a, b = 1e-9, 5e-9
    n = 200000
    z = np.linspace(a, b, n + 1)
    f = 1.0 - np.exp(-(z / lam) ** 2)
    integral = (f.sum() - 0.5 * (f[0] + f[-1])) * (b - a) / n
    assert O.mlda_pop(a, b, lam) == pytest.approx(integral / (b - a), rel=1e-8)
    # the small-cell series and the erf form agree where they meet
    v = 0.05 * lam
    assert O.mlda_pop(0.0, v * (1 - 1e-9), lam) == pytest.approx(O.mlda_pop(0.0, v * (1 + 1e-9), lam), rel=1e-6)
    # two valleys weighted: the factor lies between the single-valley ones
    m = O.silicon()
    f2 = O.mlda_factor(m, 0.5e-9, 1.5e-9, 0)
    kT = O.KB * 300.0
    l1 = O.HBAR / math.sqrt(2 * 0.916 * O.M0 * kT)
    l2 = O.HBAR / math.sqrt(2 * 0.19 * O.M0 * kT)
    lo, hi = sorted((O.mlda_pop(0.5e-9, 1.5e-9, l1), O.mlda_pop(0.5e-9, 1.5e-9, l2)))
    assert lo <= f2 <= hi


def test_entry_point_values_match_reference():
    """api3_mlda_factor of the reference for silicon at 300 and 400 K over six
    cell ranges, both carriers (tests/golden/surface_models.npz)."""
    g = load_golden("surface_models")
    for r, T in ((0, 300.0), (2, 400.0)):
        m = O.silicon(T)
        for q, (a, b) in enumerate(g["ranges"]):
            # the erf-difference form loses b/(b-a) ulps to cancellation on thin cells
            rtol = 1e-13 + 4e-16 * b / (b - a)
            got = [O.mlda_factor(m, a, b, h) for h in (0, 1)]
            np.testing.assert_allclose(got, g["mlda"][r][q], rtol=rtol, atol=0)
