"""Host logic of the measurement (bench.py): the parity criteria it reports,
the counter-summary traffic it reads, the sweep points of each workload and
the one-GPU guard.  CPU only."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402
from semiconductor_devices_simulator_amd import devices  # noqa: E402


def test_parity_against_the_reference_sweep_fixtures():
    g = np.load(os.path.join(ROOT, "tests", "golden", "e2e_mosfet_96x64x64.npz"))
    timed = bench.timed_indices(50, 5, 20)
    I = g["I"][timed].copy()
    p = bench.parity(I, timed, [96, 64, 64])
    assert p["reference"]["points_within_tolerance"] == 20 and p["reference"]["worst_err_over_bound"] == 0.0
    # the default-tolerance reference is itself within the bound of the tight one
    assert p["reference_tight"]["points_within_tolerance"] == 20
    assert p["reference_tight"]["worst_err_over_bound"] < 1.0
    # a drain current off by 1e-6 relative at a point well above its floor fails there
    F = g["F"][timed]
    j = int(np.argmax(np.abs(I[:, 1]) / (F[:, 1] + 1e-300)))
    I[j, 1] *= 1 + 1e-6
    p = bench.parity(I, timed, [96, 64, 64])
    assert p["reference"]["points_within_tolerance"] == 19
    assert p["reference"]["max_rel_drain_current_err_resolved"] == pytest.approx(1e-6, rel=1e-6)
    # one floor, not two: an error of 1.5 floors at a floor-governed current fails
    I = g["I"][timed].copy()
    k = int(np.argmin(np.abs(I[:, 1]) / (F[:, 1] + 1e-300)))
    I[k, 1] += 1.5 * F[k, 1] + 1e-8 * abs(I[k, 1])
    assert bench.parity(I, timed, [96, 64, 64])["reference"]["points_within_tolerance"] == 19
    # no fixture for the mesh: no claim
    assert bench.parity(I, timed, [7, 7, 7]) is None


def test_timed_points_span_both_drain_biases():
    spec = devices.mosfet(12, 8, 6)
    pts = bench.sweep_points(spec)
    timed = bench.timed_indices(50, 5, 20)
    assert len(timed) == 20 and timed[0] == 5 and timed[-1] == 49
    assert {pts[i][2][1] for i in timed} == {0.05, 1.0}
    text = bench.describe_timed(pts, timed, "mosfet")
    assert "Vd 0.05 V" in text and "Vd 1 V" in text
    assert bench.timed_indices(50, 2, 48) == list(range(2, 50))


def test_mixed_vs_fp64_criterion_and_fields():
    I64 = np.array([[1e-6, -1e-6], [2e-6, -2e-6]])
    F64 = np.full_like(I64, 1e-18)
    st = tuple(np.zeros(4) for _ in range(3))
    st2 = (np.full(4, 1e-10), np.zeros(4), np.zeros(4))
    out = bench.mixed_vs_fp64(I64 * (1 + 5e-9), I64, F64, st2, st)
    assert out["points_within_tolerance"] == 2
    assert out["last_point_max_abs_diff_V"]["psi"] == pytest.approx(1e-10)
    out = bench.mixed_vs_fp64(I64 * (1 + 5e-8), I64, F64, st2, st)
    assert out["points_within_tolerance"] == 0


def test_traffic_read_from_the_committed_counter_summary():
    by, src = bench.pmc_traffic([192, 128, 128])
    assert src == "profiles/k_line_scan_level0_mosfet_192x128x128_pmc.md"
    assert 40e6 < by < 160e6
    assert bench.pmc_traffic([8, 8, 8]) == (None, None)


def test_sweep_points_of_the_workloads():
    pts = bench.sweep_points(devices.mosfet(12, 8, 6))
    assert len(pts) == 50
    assert pts[0] == (2, 0.0, {1: 0.05}) and pts[25][2] == {1: 1.0}
    assert pts[24][1] == pytest.approx(1.2)
    v = devices.vdmos(8, 6, 4)
    pts = bench.sweep_points(v)
    assert len(pts) == 61 and pts[-1][0] == 4 and pts[-1][1] == pytest.approx(60.0)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def test_trace_summary_busy_time_and_window(tmp_path):
    """tracestat: the busy time is the union of the kernels' intervals (an
    overlap counts once), gaps are the idle stretches between them, and the
    window can start at a named kernel."""
    from semiconductor_devices_simulator_amd import tracestat
    p = tmp_path / "kernel_trace.csv"
    rows = [("k_setup", 0, 100), ("k_cont", 1000, 1400), ("k_a", 1300, 1500), ("k_b", 2500, 2600)]
    with open(p, "w") as f:
        f.write('"Kernel_Name","Start_Timestamp","End_Timestamp"\n')
        for n, s, e in rows:
            f.write(f'"{n}(int)",{s},{e}\n')
    S = tracestat.summarise(tracestat.load(str(p)))
    assert S["window"] == 2600 and S["busy"] == 100 + 500 + 100 and S["gaps"] == [900, 1000]
    S = tracestat.summarise(tracestat.load(str(p)), "k_cont")
    assert S["window"] == 1600 and S["busy"] == 600 and S["launches"] == 3
    assert S["per"]["k_a"] == [1, 200]
    # a counter collection: mean per dispatch, per kernel and grid size
    c = tmp_path / "counter_collection.csv"
    with open(c, "w") as f:
        f.write('"Dispatch_Id","Kernel_Name","Grid_Size","Counter_Name","Counter_Value"\n')
        for d, n, g, v in ((1, "k_a", 64, 10.0), (2, "k_a", 64, 30.0), (3, "k_a", 32, 5.0), (4, "k_b", 64, 7.0)):
            f.write(f'{d},"{n}(int)",{g},"FETCH_SIZE",{v}\n')
    P = tracestat.load_pmc(str(c))
    assert P[("k_a", "64", "FETCH_SIZE")] == (2, 20.0) and P[("k_a", "32", "FETCH_SIZE")] == (1, 5.0)
    assert list(tracestat.load_pmc(str(c), "k_b")) == [("k_b", "64", "FETCH_SIZE")]


def test_batch_sub_segments_and_worker_errors():
    """batch.sub_segment reproduces a segment's voltages piecewise, concat
    joins groups in order, and a worker that fails before touching the GPU
    (an unknown device builder) surfaces as an error in the parent."""
    from semiconductor_devices_simulator_amd import batch, devices
    spec = devices.mosfet(24, 16, 8, npts=7)
    tgt, vs, ve, n, hold = spec.sweep[1]
    V = np.linspace(vs, ve, n)
    a, b = batch.sub_segment(spec, 1, 0, 3), batch.sub_segment(spec, 1, 3)
    assert a[0] == tgt and a[4] == hold and a[3] + b[3] == n
    joined = np.concatenate([np.linspace(a[1], a[2], a[3]), np.linspace(b[1], b[2], b[3])])
    assert np.allclose(joined, V, rtol=0, atol=1e-15)
    R = batch.concat([dict(V=[1.0], I=[[1, 2]], F=[[0, 0]], conv=[True], t=[0.1], gummel=[3]),
                      dict(V=[2.0], I=[[3, 4]], F=[[0, 0]], conv=[False], t=[0.2], gummel=[5])])
    assert R["V"].tolist() == [1.0, 2.0] and R["I"].shape == (2, 2) and R["conv"].tolist() == [True, False]
    with pytest.raises(RuntimeError, match="no_such_device"):
        batch.run_concurrent(("no_such_device", (), {}), [[0]], timeout=120)
