"""The 192x128x128 n-MOSFET (3.1M nodes) on the GPU against the reference
core's full Id-Vg sweep (tests/golden/e2e_mosfet_192x128x128.npz), both
methods: all 50 bias points (Vg 0-1.2 V at Vd 0.05 V, then at Vd 1.0 V),
every terminal within 1e-8 |I_ref| + floor_ref + floor of the reference's
sweep at the same tolerances (conftest.assert_sweep_parity) and within
1e-8 |I_t| + floor_t of its sweep converged at tol_p 1e-12, tol_c 1e-11
(e2e_mosfet_192x128x128_tight.npz), and psi at the fixtures' 4096 sampled
nodes of the last point (Vd 1.0 V, Vg 1.2 V) within 1e-9 V.  The FP32 method
must have solved with its FP32 inner Krylov solves, none redone in FP64."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, assert_current_parity, assert_sweep_parity, load_golden, load_tight
from semiconductor_devices_simulator_amd import devices, sweep
from semiconductor_devices_simulator_amd.core import dptr

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.isfile(os.path.join(GOLDEN, "e2e_mosfet_192x128x128.npz")),
                    reason="full-size reference sweep fixture not generated")
@pytest.mark.parametrize("precision", [1, 0])
def test_mosfet_192x128x128_sweep_matches_reference(hip, precision):
    print(f"\n[gpu e2e] mosfet 192x128x128 precision {precision}, 50 points", flush=True)
    spec = devices.mosfet(192, 128, 128)
    g = load_golden("e2e_mosfet_192x128x128")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    assert hip.api3_solve() == 1
    hip.sdx_reset_stats()

    def progress(i, R):
        if i % 10 == 9:
            print(f"[gpu e2e]   point {i + 1} of 50", flush=True)

    R = sweep.run_sweep(hip, spec, cids, on_point=progress)
    assert len(R["I"]) == 50 and np.all(R["conv"])
    st = np.zeros(7)
    hip.sdx_kernel_stats(dptr(st), 7)
    print(f"[gpu e2e] inner solves {st[4]:.0f}, corrections {st[5]:.0f}, FP64 redone {st[6]:.0f}", flush=True)
    if precision == 1:
        assert st[4] > 0 and st[6] == 0, st
    assert_sweep_parity(R, "e2e_mosfet_192x128x128")
    t = load_tight("e2e_mosfet_192x128x128")
    assert_current_parity(R["I"], t["I"], t["F"], spec.name + " vs the tight reference sweep")
    N = spec.N
    psi, a, b = np.zeros(N), np.zeros(N), np.zeros(N)
    hip.api3_get_state(dptr(psi), dptr(a), dptr(b))
    assert np.max(np.abs(psi[g["idx"]] - g["psi"])) <= 1e-9
    assert np.max(np.abs(psi[t["idx"]] - t["psi"])) <= 1e-9
