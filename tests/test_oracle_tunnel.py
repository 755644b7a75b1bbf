"""The oracle's gate tunnelling (models bit 5) against the reference core's
at perturbed solved states of two MOSFETs with a 1.5 nm SiO2 gate stack, in
inversion and in accumulation (tests/golden/k_tun_*.npz): the paths and
supply columns, one tunnelling update (escape of every column entry,
injection of every path) and the gathered sink/source of both carriers.
Closed forms of the WKB exponent, the Simpson rule of the energy integral
and detailed balance of the fluxes."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

import semisim_oracle as O
from conftest import golden_stacks, load_golden, oracle_device

FIXTURES = ["k_tun_mosfet_12x8x6_inv", "k_tun_mosfet_12x8x6_acc", "k_tun_mosfet_20x14x9_inv"]


@pytest.fixture(scope="module", params=FIXTURES)
def case(request):
    g = load_golden(request.param)
    dev = oracle_device(g)
    P = O.tun_paths(dev, golden_stacks(g))
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    return g, dev, P, n, p


def test_paths_match_reference(case):
    g, dev, P, _, _ = case
    for key in ("pk", "pc", "pc0", "pcn", "ck"):
        np.testing.assert_array_equal(P[key], g[key], err_msg=key)
    for key in ("pA", "cl", "ch"):
        np.testing.assert_allclose(P[key], g[key], rtol=1e-15, atol=0, err_msg=key)
    # one path per gate node, every column starting at its own gate node
    assert len(P["pk"]) == int((dev.gate_mask > 0).sum())
    np.testing.assert_array_equal(P["ck"][P["pc0"]], P["pk"])


def test_update_matches_reference(case):
    g, dev, P, n, p = case
    cK, pG = O.tun_update(dev, P, g["psi"], g["phin"], g["phip"], n, p, g["Va"])
    np.testing.assert_allclose(cK, g["cK"], rtol=1e-11, atol=0)
    np.testing.assert_allclose(pG, g["pG"], rtol=1e-11, atol=0)
    for hole in (0, 1):
        tk, tg = O.tun_gather(dev, P, cK, pG, hole)
        np.testing.assert_allclose(tk, g["tk%d" % hole], rtol=1e-11, atol=0)
        np.testing.assert_allclose(tg, g["tg%d" % hole], rtol=1e-11, atol=0)
    assert np.all(cK >= 0) and np.all(pG >= 0) and pG.max() > 0


def test_wkb_closed_forms():
    m, L, c = 0.42, 1.5e-9, 2.0 * math.sqrt(2.0 * 0.42 * O.M0 * O.Q) / O.HBAR
    # a rectangular barrier: 2 kappa L
    assert O.tun_wkb(m, L, 3.1, 3.1) == pytest.approx(c * L * math.sqrt(3.1), rel=1e-14)
    # a trapezoid: the integral of sqrt over the linear profile
    for u0, u1 in ((3.1, 1.7), (0.4, 2.9), (2.0, -0.5)):
        ref = c * quad(lambda x: math.sqrt(max(u0 + (u1 - u0) * x / L, 0.0)), 0, L, epsabs=0, epsrel=1e-13,
                       limit=200)[0]
        assert O.tun_wkb(m, L, u0, u1) == pytest.approx(ref, rel=1e-9)
    # no barrier above the energy: no attenuation
    assert O.tun_wkb(m, L, -0.2, -1.0) == 0.0


def test_energy_integral_and_detailed_balance():
    ox = O.INSULATORS[8]
    D, psis, stack = O.tun_stack([(ox, 1.5e-9)], 0.3, 1.3)
    assert D == pytest.approx(1.0 / (1.5e-9 / (3.9 * O.EPS0)), rel=1e-14) and psis == 0.3
    kT, E0, EF = 0.025852, -4.3, -4.25
    # Simpson over 64 intervals against adaptive quadrature of the same integrand
    f = lambda E: O.tun_T(stack, E, 0) * math.log1p(math.exp((EF - E) / kT))
    span = min(max(EF - E0, 0.0) + 25 * kT, 3.0)
    ref = quad(f, E0, E0 + span, epsabs=0, epsrel=1e-12, limit=400)[0]
    assert O.tun_int(stack, E0, EF, kT, 0) == pytest.approx(ref, rel=1e-6)
    # holes integrate downwards with the mirrored occupation
    g = lambda E: O.tun_T(stack, E, 1) * math.log1p(math.exp(-(EF - E) / kT))
    E0h = E0 - 1.12
    span = min(max(-(EF - E0h), 0.0) + 25 * kT, 3.0)
    ref = quad(g, E0h - span, E0h, epsabs=0, epsrel=1e-12, limit=400)[0]
    assert O.tun_int(stack, E0h, EF, kT, 1) == pytest.approx(ref, rel=1e-6)
    # detailed balance: at a common Fermi level the two directions cancel
    dev = type("D", (), {"m": O.silicon()})()
    F = O.tun_flux(dev, stack, E0, E0h, EF, EF, EF)
    assert F[0] == F[1] and F[2] == F[3] and F[0] > 0 and F[2] > 0
    # a thicker oxide tunnels less, exponentially
    _, _, thick = O.tun_stack([(ox, 3.0e-9)], 0.3, 1.3)
    assert O.tun_int(thick, E0, EF, kT, 0) < 1e-3 * O.tun_int(stack, E0, EF, kT, 0)


def test_gate_current_sign_and_balance(case):
    g, dev, P, n, p = case
    cK, pG = O.tun_update(dev, P, g["psi"], g["phin"], g["phip"], n, p, g["Va"])
    I, S, In = O.tun_gate_currents(dev, P, cK, pG, n, p)
    gate = 2
    assert abs(I[gate]) > 0 and S[gate] >= abs(I[gate])
    assert np.all(I[[0, 1, 3]] == 0)


def eval_case(layers, psi, phin, phip, Vg, pm):
    si = O.silicon()
    Cref = si["chi"] + 0.5 * si["Eg"] + 0.5 * si["VT"] * math.log(si["Nc"] / si["Nv"])
    lay = [((dict(si, insulator=False) if m == 0 else O.INSULATORS[m]), t * 1e-9) for m, t in layers]
    return O.tun_eval(si, Cref, lay, psi, phin, phip, Vg, pm)


def test_tun_eval_matches_reference():
    """api3_tun_eval of the reference for six stacks (tests/golden/tun_eval_stacks.npz)."""
    import json
    g = load_golden("tun_eval_stacks")
    for c, case in enumerate(json.loads(str(g["cases"]))):
        F, J = eval_case(*case)
        np.testing.assert_allclose(F, g["F"][c], rtol=1e-11, atol=0, err_msg=str(case))
        assert J == pytest.approx(float(g["J"][c]), rel=1e-9, abs=1e-12 * max(F) * O.Q)
    # a semiconductor half cell moves the interface potential by its share of the field
    D, psis, _ = O.tun_stack([(dict(O.silicon(), insulator=False), 0.5e-9), (O.INSULATORS[8], 1.2e-9)], 0.35, 1.0)
    assert psis == pytest.approx(0.35 + D * 0.5e-9 / (11.7 * O.EPS0), rel=1e-14)
