"""Shared fixtures of the test suite.

Markers: `gpu` tests run the HIP core on a GPU (python -m pytest -m gpu);
everything else runs on the CPU (oracle against the reference's golden
vectors, host logic, the C-ABI library's exports)."""
import ctypes
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP core)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)


def golden_contacts(g):
    return json.loads(str(g["contacts"]))


def golden_stacks(g):
    """Gate stacks of a fixture: {contact index: [(material id, thickness nm)]}."""
    if "stacks" not in g:
        return {}
    return {int(k): [tuple(x) for x in v] for k, v in json.loads(str(g["stacks"])).items()}


def oracle_device(g):
    """The oracle's prepared device for a kernel fixture."""
    import semisim_oracle as O
    tau = tuple(g["tau"]) if float(g["tau"][0]) > 0 else None
    return O.Device(g["xs"], g["ys"], g["zs"], g["Nd"], g["Na"], golden_contacts(g), models=int(g["models"]), tau=tau)


@pytest.fixture(scope="session")
def hip():
    """The HIP core (ctypes, bound).  Fails loudly when it is not built."""
    from semiconductor_devices_simulator_amd import core
    return core.load_core()


def load_fixture_device(lib, g):
    """Send a kernel fixture's device through the api3 ABI (same call
    sequence as devices.load_device) and inject the fixture's state."""
    from semiconductor_devices_simulator_amd.core import dptr, iptr
    lib.api3_init()
    lib.api3_add_material(0, 300.0)
    lib.api3_set_solver(2000, 1000, 500, 1e-10, 1e-9, 1.0, 1.0)
    xs, ys, zs = (np.ascontiguousarray(g[k], np.float64) for k in ("xs", "ys", "zs"))
    assert lib.api3_set_grid_coords(len(xs), len(ys), len(zs), dptr(xs), dptr(ys), dptr(zs)) == 1
    lib.api3_set_models(int(g["models"]))
    lib.api3_set_caxis(1)
    N = len(xs) * len(ys) * len(zs)
    mat = np.zeros(N, np.int32)
    Nd, Na = (np.ascontiguousarray(g[k], np.float64) for k in ("Nd", "Na"))
    lib.api3_set_node_data(iptr(mat), dptr(Nd), dptr(Na), dptr(np.zeros(N)))
    lib.api3_clear_contacts()
    for c in golden_contacts(g):
        lib.api3_add_contact_ex(c["face"], c["i0"], c["i1"], c["j0"], c["j1"], 0, 0, c["bc"], c["V"], c["pm"],
                                c["tox"], c["qox"] / 1.602176634e-19)
    if float(g["tau"][0]) > 0:
        lib.api3_set_mat_tau(0, float(g["tau"][0]), float(g["tau"][1]))
    for ci, layers in golden_stacks(g).items():
        ids = np.asarray([lib.api3_add_material(m, 300.0) for m, _ in layers], np.int32)
        th = np.asarray([t * 1e-9 for _, t in layers], np.float64)
        assert lib.api3_set_gate_stack(ci, len(layers), iptr(ids), dptr(th)) == 1
    assert lib.sdx_test_prepare() == 1
    psi, phin, phip = (np.ascontiguousarray(g[k], np.float64) for k in ("psi", "phin", "phip"))
    lib.sdx_test_set_state(dptr(psi), dptr(phin), dptr(phip))
    return N


def current_parity(I, I_ref, F_ref):
    """Terminal-current parity against the reference: |I - I_ref| <= 1e-8
    |I_ref| + floor_ref on every terminal of every bias point, with floor_ref
    the reference's own resolution floor of that current
    (api3_get_contact_Ifloor; run_continuation, semiconductor_3d.c:4151-4171).
    Returns (ok mask, largest |I - I_ref| / (1e-8 |I_ref| + floor_ref))."""
    I, I_ref, F_ref = (np.asarray(a, np.float64) for a in (I, I_ref, F_ref))
    bound = 1e-8 * np.abs(I_ref) + F_ref
    d = np.abs(I - I_ref)
    ratio = d / np.maximum(bound, 1e-300)
    return d <= bound, float(ratio.max()) if ratio.size else 0.0


def assert_current_parity(I, I_ref, F_ref, what=""):
    ok, worst = current_parity(I, I_ref, F_ref)
    print(f"[parity] {what} worst |dI|/(1e-8|I_ref|+floor_ref) = {worst:.3g}", flush=True)
    assert ok.all(), (what, np.argwhere(~ok), np.abs(np.asarray(I) - I_ref)[~ok], np.asarray(I_ref)[~ok],
                      np.asarray(F_ref)[~ok])


def load_tight(name):
    """The reference's sweep converged far below its default tolerances
    (tol_p 1e-12, tol_c 1e-11: <name>_tight.npz), or None."""
    path = os.path.join(GOLDEN, name + "_tight.npz")
    return np.load(path, allow_pickle=False) if os.path.isfile(path) else None


def assert_sweep_parity(R, name):
    """Terminal currents of a sweep run R (sweep.run_sweep) at the default
    tolerances against the reference core's sweep of the same device at the
    same tolerances (<name>.npz): |I - I_ref| <= 1e-8 |I_ref| + floor_ref +
    floor_run on every terminal of every point.  At these tolerances each run
    pins a current only to its own floor (the Gummel loop accepts current
    changes below 3e-14 S), so two correct runs differ by up to about the sum
    of their floors: the reference against itself at 1 and 8 OpenMP threads
    differs by up to 4.65 floor_ref on the pn diode
    (tests/test_reference_reproducibility.py).  Tightening tol_c does not
    narrow this where the floor governs (the acceptance test is then the
    floor, not tol_c): runs of this core and of the reference at tol_p
    1e-12, tol_c 1e-11 still differ by up to 1.4 floors there."""
    g = load_golden(name)
    I, F = np.asarray(R["I"]), np.asarray(R["F"])
    ok, worst = current_parity(I, g["I"], g["F"] + F)
    print(f"[parity] {name} vs the reference sweep: worst |dI|/(1e-8|I_ref|+floor_ref+floor) = {worst:.3g}",
          flush=True)
    assert ok.all(), (name, np.argwhere(~ok), np.abs(I - g["I"])[~ok], g["I"][~ok], g["F"][~ok])
