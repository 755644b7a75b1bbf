"""The oracle's Anderson mixing of the Gummel map against the reference
core's aa_step (semiconductor_3d.c:3709-3757) on a solved MOSFET state: an
8-step history of a synthetic nonlinear contraction (tests/golden/
aa_mosfet_12x8x6.npz), and the closed form of the mixing on a linear map."""
import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device


@pytest.fixture(scope="module")
def case():
    g = load_golden("aa_mosfet_12x8x6")
    dev = oracle_device(g)
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    return g, dev, n, p


def check_step(k, out, g):
    """The mixed iterate of step k against the reference's: the correction
    g_out - g_in to 1e-6 of its largest entry (the Gram solve of FP32
    differences amplifies the summation order of the dots), and moved on the
    same nodes."""
    c_ref = g["Gout"][k] - g["Gin"][k]
    c = out - g["Gin"][k]
    tol = 1e-6 * np.max(np.abs(c_ref)) + 4e-15
    np.testing.assert_allclose(c, c_ref, rtol=0, atol=tol, err_msg=f"step {k}")
    assert np.array_equal(c != 0, c_ref != 0), f"step {k}: moved nodes differ"


def test_anderson_matches_reference(case):
    g, dev, n, p = case
    aa = O.Anderson(dev, n, p, int(g["aa_m"]))
    for k in range(len(g["X"])):
        out = aa.step(g["X"][k], g["Gin"][k])
        check_step(k, out, g)
    # without history the step is the plain Picard update
    assert np.array_equal(g["Gout"][0], g["Gin"][0])
    # the mixing moves psi off the contacts and the significant quasi-Fermi
    # potentials, and only those
    moved = np.zeros(3 * dev.N, bool)
    for k in range(1, len(g["X"])):
        moved |= g["Gout"][k] != g["Gin"][k]
    assert not np.any(moved & ~aa.mix)
    assert moved[: dev.N].sum() > 0 and moved[dev.N:].sum() > 0


def test_anderson_solves_a_linear_contraction(case):
    """For g(x) = xs + c (x - xs) started anywhere, the residuals of two
    iterates are parallel, so one history entry makes the mixed iterate the
    fixed point itself (secant on a linear map), up to the FP32 history and
    the 1e-9 regularisation."""
    g, dev, n, p = case
    aa = O.Anderson(dev, n, p, 4)
    rng = np.random.default_rng(3)
    xs = np.concatenate([g["psi"], g["phin"], g["phip"]])
    d = 1e-2 * rng.standard_normal(3 * dev.N) * aa.mix
    for c in (0.8, -0.5):
        aa.reset()
        x = xs + d
        g0 = xs + c * (x - xs)
        x1 = aa.step(x, g0)
        np.testing.assert_array_equal(x1, g0)              # no history yet: Picard
        x2 = aa.step(x1, xs + c * (x1 - xs))
        assert np.max(np.abs(x2 - xs)) <= 1e-6 * np.max(np.abs(d))
        # a plain Picard step would only have shrunk the error by c again
        assert np.max(np.abs(xs + c * c * d - xs)) > 0.2 * np.max(np.abs(d))


def test_history_is_a_ring_of_depth_m(case):
    g, dev, n, p = case
    aa = O.Anderson(dev, n, p, 2)
    for k in range(5):
        aa.step(g["X"][k], g["Gin"][k])
    assert aa.n == 2 and aa.pos == 0          # 4 differences into 2 slots
    np.testing.assert_allclose(aa.gram[0, 1], aa.gram[1, 0], rtol=0, atol=0)
