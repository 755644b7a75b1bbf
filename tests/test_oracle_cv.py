"""The oracle's terminal partition and terminal charges (quasi-static C-V)
against the reference core's at solved MOSFET states, on and off
(tests/golden/cv_*.npz), plus charge neutrality."""
import numpy as np
import pytest

import semisim_oracle as O
from conftest import load_golden, oracle_device


@pytest.mark.parametrize("case", ["on", "off"])
def test_partition_and_charges_match_reference(case):
    g = load_golden(f"cv_mosfet_24x16x8_{case}")
    dev = oracle_device(g)
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    own, meet = O.charge_partition(dev, n, p, list(g["cterm"]))
    np.testing.assert_array_equal(own, g["owner"])
    assert meet == int(g["meet"])
    Q = O.terminal_charges(dev, g["psi"], own, list(g["cterm"]), g["Va"])
    scale = np.max(np.abs(g["Q0"]))
    assert np.all(np.abs(Q - g["Q0"]) <= 1e-12 * scale), (Q, g["Q0"])
    # Gauss's law over the whole device: every node has a terminal, so the
    # fluxes between terminals cancel pairwise and the charges sum to about 0
    assert np.all(own >= 0)
    assert abs(Q[:-1].sum()) <= 1e-6 * scale
