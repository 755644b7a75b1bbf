"""End-to-end bias sweeps on the GPU against the reference core's sweeps
(golden fixtures written by tests/golden/make_golden.py).  Parity criteria
per bias point (conftest.assert_sweep_parity): |dI| <= 1e-8 |I_t| + floor_t
against the reference's sweep converged far below its default tolerances,
and |dI| <= 1e-8 |I_ref| + floor_ref + floor_hip against its sweep at the same
tolerances (each run resolves a current only to its own floor); |d psi| <=
1e-9 V at the last point."""
import numpy as np
import pytest

from conftest import assert_current_parity, assert_sweep_parity, load_golden, load_tight
from semiconductor_devices_simulator_amd import devices, sweep
from semiconductor_devices_simulator_amd.core import dptr

pytestmark = pytest.mark.gpu


def run(hip, spec, precision=0):
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    assert hip.api3_solve() == 1
    return cids, sweep.run_sweep(hip, spec, cids)


def check_currents(R, name):
    assert np.all(R["conv"])
    assert_sweep_parity(R, name)


def state(hip, N):
    a, b, c = np.zeros(N), np.zeros(N), np.zeros(N)
    hip.api3_get_state(dptr(a), dptr(b), dptr(c))
    return a, b, c


@pytest.mark.parametrize("precision", [0, 1])
def test_pn_diode_sweep_matches_reference(hip, precision):
    print(f"\n[gpu e2e] pn diode 16x16x64 precision {precision}", flush=True)
    spec = devices.pn_diode()
    g = load_golden("e2e_pn_diode_16x16x64")
    _, R = run(hip, spec, precision)
    check_currents(R, "e2e_pn_diode_16x16x64")
    psi, phin, phip = state(hip, spec.N)
    assert np.max(np.abs(psi - g["psi"])) <= 1e-9
    assert np.max(np.abs(phin - g["phin"])) <= 1e-8
    # Kirchhoff: the raw contact fluxes sum to zero within the floors
    Iraw = [hip.api3_get_contact_Iraw(c) for c in range(2)]
    assert abs(sum(Iraw)) <= 4 * max(hip.api3_get_contact_Ifloor(c) for c in range(2)) + 1e-8 * abs(Iraw[0])


@pytest.mark.parametrize("precision", [0, 1])
def test_small_mosfet_sweep_matches_reference(hip, precision):
    print(f"\n[gpu e2e] mosfet 24x16x8 precision {precision}", flush=True)
    spec = devices.mosfet(24, 16, 8, npts=7)
    g = load_golden("e2e_mosfet_24x16x8")
    _, R = run(hip, spec, precision)
    check_currents(R, "e2e_mosfet_24x16x8")
    psi, _, _ = state(hip, spec.N)
    assert np.max(np.abs(psi - g["psi"])) <= 1e-9


@pytest.mark.parametrize("precision", [0, 1])
def test_shaped_contact_sweep_matches_reference(hip, precision):
    """A shaped box contact sent as a node list (api3_set_con_nodes, as the
    front end sends shaped contacts): the pn diode with its anode a two-plane
    staircase disc, forward sweep, against the reference's sweep of the same
    device; a node list with a node outside the mesh is refused."""
    from semiconductor_devices_simulator_amd.core import iptr
    print(f"\n[gpu e2e] shaped anode diode 16x16x64 precision {precision}", flush=True)
    spec = devices.shaped_anode_diode()
    g = load_golden("e2e_shaped_anode_diode_16x16x64")
    cids, R = run(hip, spec, precision)
    check_currents(R, "e2e_shaped_anode_diode_16x16x64")
    psi, _, _ = state(hip, spec.N)
    assert np.max(np.abs(psi - g["psi"])) <= 1e-9
    # the full-face diode differs: the list is what was contacted
    full = load_golden("e2e_pn_diode_16x16x64")
    assert np.max(np.abs(R["I"][-1, 0] - full["I"][-1, 0])) > 1e-3 * abs(full["I"][-1, 0])
    bad = np.array([0, spec.N], np.int32)
    assert hip.api3_set_con_nodes(cids[0], 2, iptr(bad), None) == 0


def test_runs_are_bit_reproducible(hip):
    print("\n[gpu e2e] reproducibility", flush=True)
    spec = devices.mosfet(24, 16, 8, npts=3)
    _, R1 = run(hip, spec)
    p1 = state(hip, spec.N)[0]
    _, R2 = run(hip, spec)
    p2 = state(hip, spec.N)[0]
    assert np.array_equal(R1["I"], R2["I"]) and np.array_equal(p1, p2)


def test_stop_request_keeps_last_state(hip):
    spec = devices.pn_diode(8, 8, 32)
    cids = devices.load_device(hip, spec)
    assert hip.api3_solve() == 1
    before = state(hip, spec.N)[0]
    hip.api3_set_contact_V(cids[0], 0.5)
    hip.api3_request_stop()
    try:
        assert hip.api3_resolve() == 0
        assert hip.api3_stop_requested() == 1
    finally:
        hip.api3_clear_stop()
    np.testing.assert_array_equal(state(hip, spec.N)[0], before)
    assert hip.api3_get_contact_Vsolved(cids[0]) == 0.0
    assert hip.api3_resolve() == 1
    assert hip.api3_get_contact_Vsolved(cids[0]) == 0.5


def test_state_injection_round_trip(hip):
    spec = devices.pn_diode(8, 8, 32)
    cids = devices.load_device(hip, spec)
    hip.api3_set_contact_V(cids[0], 0.3)
    assert hip.api3_solve() == 1
    s = state(hip, spec.N)
    I0 = hip.api3_get_contact_I(cids[0])
    assert hip.api3_set_state(dptr(s[0]), dptr(s[1]), dptr(s[2])) == 1
    assert hip.api3_resolve() == 1                  # no bias change: one Gummel solve from the state
    assert hip.api3_get_contact_I(cids[0]) == pytest.approx(I0, rel=1e-8)
    assert hip.api3_get_iter() <= 3


def test_out_of_scope_physics_is_rejected(hip):
    spec = devices.pn_diode(8, 8, 32)
    spec.contacts[1].bc = 1                          # Schottky cathode
    devices.load_device(hip, spec)
    hip.api3_set_models(64)                          # thermionic emission at a Schottky face contact
    assert hip.api3_solve() == 0
    hip.api3_set_models(0)
    assert hip.api3_solve() == 1


def test_readback_quantities(hip):
    spec = devices.pn_diode(8, 8, 32)
    cids = devices.load_device(hip, spec)
    hip.api3_set_contact_V(cids[0], 0.6)
    assert hip.api3_solve() == 1
    N = spec.N
    out = np.zeros(N)
    hip.api3_bulk_copy(13, dptr(out))                 # kind: all semiconductor or contact metal-free
    assert set(np.unique(out)) <= {0.0}
    hip.api3_bulk_copy(18, dptr(out))                 # contact id per node
    assert out[0] == 1.0 and out[-1] == 2.0
    hip.api3_bulk_copy(9, dptr(out))                  # Jz: current flows along z
    Jz = out.copy()
    A = hip.api3_get_contact_A(cids[0])
    I = hip.api3_get_contact_I(cids[0])
    mid = Jz.reshape(32, 8, 8)[16]
    assert np.mean(mid) * A == pytest.approx(I, rel=0.05)        # anode current flows in +z
    n, p = np.zeros(N), np.zeros(N)
    hip.api3_bulk_copy(1, dptr(n))
    hip.api3_bulk_copy(2, dptr(p))
    assert np.all(n > 0) and np.all(p > 0)


@pytest.mark.parametrize("precision", [0, 1])
def test_c_side_sweeps_match_reference(hip, precision):
    """api3_iv_sweep2 (all currents and floors per point) of the gate of the
    24x16x8 MOSFET from equilibrium at Vd = 0.05 V, then api3_iv_sweep
    (current density of the swept contact) of the drain from 0.05 to 1.0 V
    at Vg = 1.2 V, against the reference core's same two calls."""
    from semiconductor_devices_simulator_amd.core import iptr
    print(f"\n[gpu e2e] C-side sweeps mosfet 24x16x8 precision {precision}", flush=True)
    spec = devices.mosfet(24, 16, 8, npts=7)
    g = load_golden("e2e_ivsweep_mosfet_24x16x8")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    assert hip.api3_solve() == 1
    hip.api3_set_contact_V(cids[1], 0.05)
    n, nc = 7, len(cids)
    V, I, F, C = np.zeros(n), np.zeros(n * nc), np.zeros(n * nc), np.zeros(n, np.int32)
    hip.api3_iv_sweep2(cids[2], 0.0, 1.2, n, dptr(V), dptr(I), dptr(F), iptr(C))
    I, F = I.reshape(n, nc), F.reshape(n, nc)
    assert np.all(C == 1) and np.all(g["C"] == 1)
    np.testing.assert_allclose(V, g["V"], rtol=0, atol=1e-15)
    assert_current_parity(I, g["I"], g["F"] + F, "iv_sweep2 vs the reference's iv_sweep2")
    t = load_tight("e2e_mosfet_24x16x8")          # the same gate sweep, converged far tighter
    assert_current_parity(I, t["I"][:n], t["F"][:n], "iv_sweep2 vs the tight reference sweep")
    V1, J1, C1 = np.zeros(n), np.zeros(n), np.zeros(n, np.int32)
    hip.api3_set_contact_V(cids[2], 1.2)
    hip.api3_iv_sweep(cids[1], 0.05, 1.0, n, dptr(V1), dptr(J1), iptr(C1))
    assert np.all(C1 == 1) and np.all(g["C1"] == 1)
    np.testing.assert_allclose(V1, g["V1"], rtol=0, atol=1e-15)
    # J = I/A of the swept contact (drain), compared at 1e-8 relative plus
    # the largest current floor of the drain in the first sweep, scaled by A
    A = g["A"][1]
    assert hip.api3_get_contact_A(cids[1]) == pytest.approx(A, rel=1e-14)
    assert np.all(np.abs(g["J1"]) > 0)
    Fl = (g["F"][:, 1].max() + F[:, 1].max()) / A
    assert np.all(np.abs(J1 - g["J1"]) <= 1e-8 * np.abs(g["J1"]) + Fl), (J1, g["J1"])


def test_unit_diagonal_products_are_bitwise(hip):
    """The continuity BiCGSTAB products skip the stored unit diagonal (tune
    20); multiplying by 1 is exact, so a sweep's currents are bit-identical
    with the diagonal read."""
    print("\n[gpu e2e] unit-diagonal products, pn diode", flush=True)
    spec = devices.pn_diode()
    seg = [(0, 0.0, 0.3, 4, {})]
    out = []
    try:
        for ud in (1, 0):
            cids = devices.load_device(hip, spec)
            hip.sdx_tune(20, ud)
            assert hip.sdx_set_precision(1) == 1
            assert hip.api3_solve() == 1
            R = sweep.run_sweep(hip, spec, cids, segments=seg)
            assert np.all(R["conv"])
            out.append((R["I"], state(hip, spec.N)))
    finally:
        hip.sdx_tune(20, 1)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [0, 1])
def test_staged_reads_are_bitwise(hip, precision):
    """The Krylov updates read the V-cycle results where the V-cycles left
    them (two level-0 arrays, tune 23) instead of FP64 copies written by the
    products: the same values, so a sweep's currents and fields are
    bit-identical."""
    print(f"\n[gpu e2e] staged reads precision {precision}, mosfet 24x16x8", flush=True)
    spec = devices.mosfet(24, 16, 8, npts=4)
    seg = spec.sweep[:1]
    out = []
    try:
        for on in (1, 0):
            cids = devices.load_device(hip, spec)
            hip.sdx_tune(23, on)
            assert hip.sdx_set_precision(precision) == 1
            assert hip.api3_solve() == 1
            R = sweep.run_sweep(hip, spec, cids, segments=seg)
            assert np.all(R["conv"])
            out.append((R["I"], state(hip, spec.N), [hip.api3_get_stat(i) for i in range(4)]))
    finally:
        hip.sdx_tune(23, 1)
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("precision", [0, 1])
def test_readback_matches_reference(hip, precision):
    """All 36 api3_bulk_copy quantities and api3_face_J of the 24x16x8
    MOSFET solved at Vd = 0.5 V, Vg = 1.0 V against the reference core's at
    the same bias (rb_mosfet_24x16x8.npz).  Both cores converge to the same
    state within the solver tolerance (psi to 1e-9 V), so quantities that
    follow from the state match to the tolerances below; the integer and
    input-derived ones (kind, doping, contact ids, interface distance, the
    quantities of physics outside the scope) match exactly."""
    print(f"\n[gpu e2e] readback parity precision {precision}", flush=True)
    spec = devices.mosfet(24, 16, 8)
    g = load_golden("rb_mosfet_24x16x8")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    for c, v in enumerate(g["Va"]):
        hip.api3_set_contact_V(cids[c], float(v))
    assert hip.api3_solve() == 1
    N = spec.N
    bulk = np.zeros((36, N))
    for w in range(36):
        hip.api3_bulk_copy(w, dptr(bulk[w]))
    ref = g["bulk"]

    def close(w, rtol, atol_frac=0.0, atol=0.0, scale=None):
        a, b = bulk[w], ref[w]
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=f"quantity {w}: NaN pattern")
        m = ~np.isnan(b)
        sc = np.max(np.abs(b[m])) if scale is None else scale
        tol = rtol * np.abs(b[m]) + atol_frac * sc + atol
        err = np.abs(a[m] - b[m])
        assert np.all(err <= tol), (w, float(err.max()), float(np.max(err - tol)))

    close(0, 0, atol=1e-9)                                  # psi (V)
    for w in (1, 2):                                        # n, p
        close(w, 1e-6)
    for w in (3, 4):                                        # band edges (eV)
        close(w, 0, atol=1e-9)
    for w in (5, 6, 15, 16):                                # quasi-Fermi levels
        close(w, 0, atol=1e-6)
    # current densities: 1e-4 relative plus 1e-6 of the largest total current
    # density.  The solves resolve currents relative to the total current, so
    # a component across the current flow (Jz of this device) and the
    # minority (hole) current, 1e-16 of the electron current here, compare
    # only at that scale
    sc = max(np.max(np.abs(ref[w])) for w in (7, 8, 9))
    for w in (7, 8, 9, 20, 21, 22, 23, 24, 25):
        close(w, 1e-4, atol_frac=1e-6, scale=sc)
    for w in (10, 11):                                      # mobilities
        close(w, 1e-6)
    close(12, 1e-4, atol_frac=1e-6)                         # recombination
    for w in (13, 14, 17, 18, 27, 28, 29, 31, 32, 33, 34, 35):
        close(w, 0)                                         # exact
    for w in (19, 26, 30):                                  # BGN, n_i, interface distance
        close(w, 1e-12)
    sc = np.max(np.abs(g["faceJ"][:, 0]))
    for car in range(2):
        for ax in range(3):
            fj = np.zeros(N)
            assert hip.api3_face_J(ax, car, dptr(fj)) == 1
            r = g["faceJ"][ax, car]
            assert np.all(np.abs(fj - r) <= 1e-4 * np.abs(r) + 1e-6 * sc), (ax, car)


@pytest.mark.parametrize("case", ["on", "off"])
@pytest.mark.parametrize("precision", [0, 1])
def test_terminal_charges_match_reference(hip, precision, case):
    """Quasi-static C-V support (api3_charge_partition, api3_charge_owner,
    api3_terminal_charges) on the 24x16x8 MOSFET at Vd 0.5 / Vg 1.0 V (on)
    and at 0 V (off), and the gate column of the capacitance matrix by
    central differences (+-5 mV, one resolve each) as the front end's
    cap_matrix forms it, against the reference core's same calls."""
    from semiconductor_devices_simulator_amd.core import iptr
    print(f"\n[gpu e2e] terminal charges {case} precision {precision}", flush=True)
    g = load_golden(f"cv_mosfet_24x16x8_{case}")
    spec = devices.mosfet(24, 16, 8)
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    Va = [float(v) for v in g["Va"]]
    for c, v in enumerate(Va):
        hip.api3_set_contact_V(cids[c], v)
    assert hip.api3_solve() == 1
    N, ct = spec.N, np.ascontiguousarray(g["cterm"], np.int32)
    nt = int(ct.max()) + 1
    meet = hip.api3_charge_partition(iptr(ct))
    own = np.zeros(N, np.int32)
    assert hip.api3_charge_owner(iptr(own)) == 1
    # the partition follows the majority carrier (n >= p) of every node; the
    # two states agree to the solver tolerance, so it is the same node set
    np.testing.assert_array_equal(own, g["owner"])
    assert meet == int(g["meet"])
    Q0 = np.zeros(nt + 1)
    assert hip.api3_terminal_charges(nt, dptr(Q0)) == 1
    scale = np.max(np.abs(g["Q0"]))
    assert np.all(np.abs(Q0 - g["Q0"]) <= 1e-6 * scale), (Q0, g["Q0"])
    tgt, dV = int(g["tgt"]), float(g["dV"])
    Qs = []
    for sg in (1.0, -1.0):
        hip.api3_set_contact_V(cids[tgt], Va[tgt] + sg * dV)
        assert hip.api3_resolve() == 1
        Q = np.zeros(nt + 1)
        assert hip.api3_terminal_charges(nt, dptr(Q)) == 1
        Qs.append(Q)
    K, Kr = (Qs[0] - Qs[1]) / (2 * dV), (g["Qp"] - g["Qm"]) / (2 * dV)
    assert np.all(np.abs(K - Kr) <= 1e-4 * np.max(np.abs(Kr))), (K, Kr)
    # charge neutrality of the column: the terminal charges' changes sum to 0
    assert abs(K[:nt].sum()) <= 1e-3 * np.max(np.abs(K))


@pytest.mark.parametrize("precision", [0, 1])
def test_scalar_readback_matches_reference(hip, precision):
    """Scalar readbacks at the same solved bias point as the node readback
    (rb_mosfet_24x16x8.npz): every contact's current, electron current, raw
    flux, floor and area; the carrier numbers; the largest semiconductor
    field; the largest insulator (gate oxide) field and where it is."""
    print(f"\n[gpu e2e] scalar readback parity precision {precision}", flush=True)
    spec = devices.mosfet(24, 16, 8)
    g = load_golden("rb_mosfet_24x16x8")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    for c, v in enumerate(g["Va"]):
        hip.api3_set_contact_V(cids[c], float(v))
    assert hip.api3_solve() == 1
    ref = g["contacts_IInIrawFA"]
    for c in range(len(cids)):
        I, In, Iraw = (hip.api3_get_contact_I(cids[c]), hip.api3_get_contact_In(cids[c]),
                       hip.api3_get_contact_Iraw(cids[c]))
        F, A = hip.api3_get_contact_Ifloor(cids[c]), hip.api3_get_contact_A(cids[c])
        fl = F + ref[c, 3]
        for got, want in ((I, ref[c, 0]), (In, ref[c, 1]), (Iraw, ref[c, 2])):
            assert abs(got - want) <= 1e-8 * abs(want) + fl, (c, got, want)
        assert F == pytest.approx(ref[c, 3], rel=1e-3, abs=1e-30)
        assert A == pytest.approx(ref[c, 4], rel=1e-14)
    sums = np.zeros(2)
    hip.api3_carrier_sums(dptr(sums))
    np.testing.assert_allclose(sums, g["carrier_sums"], rtol=1e-6)
    assert hip.api3_get_Emax() == pytest.approx(float(g["Emax"]), rel=1e-6)
    pos = np.zeros(3)
    assert hip.api3_ins_field_max(dptr(pos)) == pytest.approx(float(g["ins_field_max"]), rel=1e-6)
    # the device is uniform along z (the width), so the oxide field takes its
    # largest value at every z of one (x, y) to rounding: the position must
    # agree in x and y, and z is any mesh plane
    np.testing.assert_allclose(pos[:2], g["ins_field_pos"][:2], rtol=0, atol=1e-15)
    assert np.min(np.abs(spec.zs - pos[2])) <= 1e-15
