"""Gate tunnelling (models bit 5) on the device against the reference core.
Kernel level, on the k_tun_* fixtures: the paths and supply columns prepare
builds (exact), one tunnelling update (escape of every column entry,
injection of every path) and the gathered sink/source of both carriers
against the reference's and the oracle's.  End to end: the Id-Vg sweep of a
24x16x8 MOSFET with a 1.5 nm SiO2 gate stack, every terminal (the gate
current included) against the reference's sweep, both methods."""
import ctypes

import numpy as np
import pytest

import semisim_oracle as O
from conftest import assert_sweep_parity, golden_stacks, load_golden, load_fixture_device, oracle_device
from semiconductor_devices_simulator_amd import devices, sweep
from semiconductor_devices_simulator_amd.core import dptr, iptr

pytestmark = pytest.mark.gpu

FIXTURES = ["k_tun_mosfet_12x8x6_inv", "k_tun_mosfet_12x8x6_acc", "k_tun_mosfet_20x14x9_inv"]


@pytest.mark.parametrize("name", FIXTURES)
def test_tunnelling_update_matches_reference(hip, name):
    print(f"\n[gpu tunnel] fixture {name}", flush=True)
    g = load_golden(name)
    N = load_fixture_device(hip, g)
    ncol = ctypes.c_int(0)
    npth = hip.sdx_test_tun_paths(ctypes.byref(ncol), None, None, None, None, None, None, None, None)
    assert npth == len(g["pk"]) and ncol.value == len(g["ck"])
    assert hip.api3_get_ntp(-1) == npth and hip.api3_get_ntp(2) == npth
    pk, pc, pc0, pcn = (np.zeros(npth, np.int32) for _ in range(4))
    pA = np.zeros(npth)
    ck = np.zeros(ncol.value, np.int32)
    cl, ch = np.zeros(ncol.value), np.zeros(ncol.value)
    hip.sdx_test_tun_paths(ctypes.byref(ncol), iptr(pk), iptr(pc), iptr(pc0), iptr(pcn), dptr(pA), iptr(ck),
                           dptr(cl), dptr(ch))
    for a, key in ((pk, "pk"), (pc, "pc"), (pc0, "pc0"), (pcn, "pcn"), (ck, "ck")):
        np.testing.assert_array_equal(a, g[key], err_msg=key)
    for a, key in ((pA, "pA"), (cl, "cl"), (ch, "ch")):
        np.testing.assert_allclose(a, g[key], rtol=1e-15, atol=0, err_msg=key)
    Va = np.ascontiguousarray(g["Va"], np.float64)
    dev = oracle_device(g)
    P = O.tun_paths(dev, golden_stacks(g))
    n, p = dev.carriers(g["psi"], g["phin"], g["phip"])
    oK, oG = O.tun_update(dev, P, g["psi"], g["phin"], g["phip"], n, p, g["Va"])
    for hole in (0, 1):
        cK, pG, tk, tg = np.zeros(2 * ncol.value), np.zeros(2 * npth), np.zeros(N), np.zeros(N)
        assert hip.sdx_test_tun(dptr(Va), hole, dptr(cK), dptr(pG), dptr(tk), dptr(tg)) == npth
        np.testing.assert_allclose(cK, g["cK"], rtol=1e-11, atol=0)
        np.testing.assert_allclose(pG, g["pG"], rtol=1e-11, atol=0)
        np.testing.assert_allclose(cK, oK, rtol=1e-11, atol=0)
        np.testing.assert_allclose(pG, oG, rtol=1e-11, atol=0)
        np.testing.assert_allclose(tk, g["tk%d" % hole], rtol=1e-11, atol=0)
        np.testing.assert_allclose(tg, g["tg%d" % hole], rtol=1e-11, atol=0)


@pytest.mark.parametrize("precision", [0, 1])
def test_tunnelling_sweep_matches_reference(hip, precision):
    print(f"\n[gpu tunnel] mosfet 24x16x8 models 39 tox 1.5 nm precision {precision}", flush=True)
    spec = devices.with_gate_tunnelling(devices.mosfet(24, 16, 8, npts=7, tox=1.5))
    g = load_golden("e2e_tun_mosfet_24x16x8")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    assert hip.api3_solve() == 1
    assert hip.api3_get_ntp(2) > 0
    R = sweep.run_sweep(hip, spec, cids)
    assert np.all(R["conv"])
    assert_sweep_parity(R, "e2e_tun_mosfet_24x16x8")
    assert np.all(np.abs(R["I"][:, 2]) > 1e-13)               # a gate current, resolved
    psi = np.zeros(spec.N)
    hip.api3_get_state(dptr(psi), dptr(np.zeros(spec.N)), dptr(np.zeros(spec.N)))
    assert np.max(np.abs(psi - g["psi"])) <= 1e-9
    # the gate current density readback integrates to the gate current
    J = np.zeros(spec.N)
    hip.api3_bulk_copy(27, dptr(J))
    dev_area = np.zeros(spec.N)
    dev = O.Device(spec.xs, spec.ys, spec.zs, spec.Nd, spec.Na,
                   [dict(face=c.face, i0=i0, i1=i1, j0=j0, j1=j1, bc=c.bc, V=c.V, pm=c.pm, tox=c.tox * 1e-9,
                         qox=0.0) for c, (i0, i1, j0, j1, _, _) in
                    ((c, c.to_idx6(spec.xs, spec.ys, spec.zs)) for c in spec.contacts)], models=spec.models)
    dev_area[:] = dev.robA
    assert np.sum(J * dev_area) == pytest.approx(R["I"][-1, 2], rel=1e-9)


def test_tun_eval_matches_reference(hip):
    """api3_tun_eval (fluxes through a given stack, evaluated by the device
    kernels of the solve) against the reference's for six stacks."""
    import json
    print("\n[gpu tunnel] api3_tun_eval, 6 stacks", flush=True)
    g = load_golden("tun_eval_stacks")
    hip.api3_init()
    ids = {m: hip.api3_add_material(m, 300.0) for m in (0, 8, 10, 11)}
    for c, (layers, psi, phin, phip, Vg, pm) in enumerate(json.loads(str(g["cases"]))):
        mi = np.asarray([ids[m] for m, _ in layers], np.int32)
        th = np.asarray([t * 1e-9 for _, t in layers], np.float64)
        out = np.zeros(4)
        J = hip.api3_tun_eval(ids[0], len(layers), iptr(mi), dptr(th), psi, phin, phip, Vg, pm, dptr(out))
        np.testing.assert_allclose(out, g["F"][c], rtol=1e-11, atol=0, err_msg=str(layers))
        assert J == pytest.approx(float(g["J"][c]), rel=1e-9, abs=1e-12 * max(out) * 1.602176634e-19)
    # invalid arguments return 0 and leave the output alone, as the reference
    out = np.full(4, 7.0)
    assert hip.api3_tun_eval(ids[0], 0, iptr(mi), dptr(th), 0.0, 0.0, 0.0, 1.0, 4.65, dptr(out)) == 0.0
    assert np.all(out == 7.0)


@pytest.mark.parametrize("precision", [0, 1])
def test_meshed_oxide_tunnelling_sweep_matches_reference(hip, precision):
    """Gate tunnelling through meshed insulator nodes (tun_build :1214-1237):
    the diode of devices.meshed_gate_diode, whose planes x = 0 and 1 nm are
    SiO2 under a metal gate, builds the same 176 paths as the reference (a
    silicon half cell, then the oxide to the metal) and none without the gate;
    its gate sweep 0 -> 2 V (models 7|32) matches the reference's sweep at every
    terminal, the gate's tunnelling current included (up to 1.8e-8 A), within
    1e-8 |I_ref| + floor_ref + floor, and psi of the last point within 1e-9 V."""
    print(f"\n[gpu tunnel] meshed gate diode precision {precision}", flush=True)
    spec = devices.meshed_gate_diode()
    g = load_golden("e2e_meshed_gate_diode_8x8x32")
    cids = devices.load_device(hip, spec)
    assert hip.sdx_set_precision(precision) == 1
    assert hip.api3_solve() == 1
    ntp = [hip.api3_get_ntp(-1)] + [hip.api3_get_ntp(c) for c in range(len(cids))]
    assert ntp == g["ntp"].tolist(), (ntp, g["ntp"])
    R = sweep.run_sweep(hip, spec, cids)
    assert np.all(R["conv"])
    print("[gpu tunnel] gate current", R["I"][:, 2].tolist(), flush=True)
    assert_sweep_parity(R, "e2e_meshed_gate_diode_8x8x32")
    N = spec.N
    psi = np.zeros(N)
    hip.api3_get_state(dptr(psi), dptr(np.zeros(N)), dptr(np.zeros(N)))
    assert np.max(np.abs(psi - g["psi"])) <= 1e-9
    # without the gate the march ends on no metal: no path
    nogate = devices.meshed_gate_diode()
    nogate.contacts = nogate.contacts[:2]
    nogate.sweep = [(0, 0.0, 0.0, 1, {})]
    devices.load_device(hip, nogate)
    assert hip.sdx_test_prepare() == 1 and hip.api3_get_ntp(-1) == 0
