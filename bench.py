#!/usr/bin/env python3
"""Bias points per second of the n-MOSFET Id-Vg sweep on one MI355X.

Workload: the planar n-MOSFET on a 192x128x128 tensor mesh (3,145,728 nodes),
Id-Vg from 0 to 1.2 V in 0.05 V steps at Vd = 0.05 V and 1.0 V (50 bias
points), synthetic device (no external data).  A step is one bias point:
api3_resolve continuing the sweep from the previous point, converged to the
reference's tolerances (tol_p 1e-10 V, tol_c 1e-9 relative current), inputs
already resident on the device.  K timed points are spread evenly over sweep
indices W..49, so both drain biases are timed; the points between them are
solved untimed (the continuation needs them).  config.workload and
config.timed_points name exactly which points were timed.  --mesh 96 64 64
runs the 393,216-node mesh instead.  --workload vdmos runs the 55 V VDMOS
cell with impact ionisation (256x128x256, 8,388,608 nodes; drain 0 -> 60 V in
1 V steps at Vgs = 0) as a supplementary line: the headline metric is quoted
on the MOSFET.  --workload prism runs the graph (prism-mesh) path: the planar
n-MOSFET on a triangular-prism mesh (257 x 129 lattice x 16 layers, 531,000
nodes; Id-Vg 0 -> 1.2 V in 0.1 V steps at Vd = 0.05 V), FP64 aggregation
multigrid, a supplementary line whose roofline is the graph Krylov SpMV.

Up to three solver configurations run back to back on the same sweep: the
plain FP64 method (FP64 multigrid inside FP64 PCG/BiCGSTAB at the reference's
tolerances), the mixed-precision method (FP32 multigrid; each linear solve is
FP64 defect correction around an FP32 Krylov solve whose products run on the
FP32 level-0 operator, with FP64 residuals and FP64-accumulated dots; its own
inner tolerances: Newton-step PCG 1e-2, continuity BiCGSTAB 1e-3 while the
Gummel update exceeds 1e-6 V and the reference's 1e-4 after), and the
mixed-precision method seeded by the learned warm start when the trained
network (semiconductor_devices_simulator_amd/warmstart_unet64.npz) is present.
`value` is the learned method's rate when its warm start kept corrections and
beat the mixed-precision method alone, else the mixed-precision rate; every
rate and the speedup over plain FP64 are reported beside it.  Parity: the
terminal currents of every timed point of each method within 1e-8 |I_ref| +
floor_ref of the reference core's sweep of the same mesh
(tests/golden/e2e_mosfet_<mesh>.npz) and of its sweep converged at tol_p
1e-12, tol_c 1e-11 (..._tight.npz) when present, and the mixed-precision
currents and final fields against the plain FP64 GPU run.

roofline: the dominant kernel (the level-0 zebra line Gauss-Seidel sweep of
the multigrid smoother), its algorithmic bytes per launch divided by its
average launch time, measured with HIP events on the solver stream during the
timed region (every 8th V-cycle runs with its level-0 sweeps bracketed).
cpu_baseline: the reference C core (oracle/_ref/libsemisim_ref.so, built from
the reference sources) on the host cores allotted to the GPU, on a sample of
the same sweep spanning both drain biases: at each sample position it resumes
from the GPU's converged state, solves one point untimed and is timed on the
next.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--mesh NX NY NZ] [--workload mosfet|vdmos|prism]
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SWEEP = {"mosfet": "Id-Vg sweep (0-1.2 V x 0.05 V at Vd 0.05 and 1.0 V)",
         "vdmos": "drain sweep (0-60 V x 1 V at Vgs 0, impact ionisation)",
         "prism": "Id-Vg sweep (0-1.2 V x 0.1 V at Vd 0.05 V) on a triangular-prism mesh"}
DEFAULT_MESH = {"mosfet": [192, 128, 128], "vdmos": [256, 128, 256], "prism": [257, 129, 16]}
DRAIN = {"mosfet": 1, "vdmos": 4, "prism": 1, "prism_mosfet": 1}   # contact index of the drain
PRISM_NPTS = 13


def make_spec(workload, mesh):
    from semiconductor_devices_simulator_amd import devices, prism
    if workload == "prism":
        return prism.prism_mosfet(*mesh, npts=PRISM_NPTS)
    return devices.mosfet(*mesh) if workload == "mosfet" else devices.vdmos(*mesh)


def load_spec(lib, spec):
    from semiconductor_devices_simulator_amd import devices, prism
    return (prism.load_graph_device if isinstance(spec, prism.GraphSpec) else devices.load_device)(lib, spec)


def sweep_points(spec):
    pts = []
    for tgt, vs, ve, n, hold in spec.sweep:
        for v in np.linspace(vs, ve, n):
            pts.append((tgt, float(v), hold))
    return pts


def solve_points(lib, cids, pts, i0, i1, floors=None):
    """Solve sweep points i0..i1-1; returns their terminal currents (and
    appends the current floors to `floors` when given)."""
    I = []
    for tgt, v, hold in pts[i0:i1]:
        for c, hv in hold.items():
            lib.api3_set_contact_V(cids[c], hv)
        lib.api3_set_contact_V(cids[tgt], v)
        ok = lib.api3_resolve()
        if not ok:
            raise RuntimeError(f"bias point V={v} did not converge")
        I.append([lib.api3_get_contact_I(cids[c]) for c in range(len(cids))])
        if floors is not None:
            floors.append([lib.api3_get_contact_Ifloor(cids[c]) for c in range(len(cids))])
    return np.array(I)


def timed_indices(npts, warmup, steps):
    """The bias points that are timed: `steps` indices spread evenly over
    [warmup, npts), so that every segment of the sweep (both drain biases of
    the Id-Vg sweep) is represented; the points between them are solved
    untimed, as the continuation needs them."""
    idx = np.unique(np.round(np.linspace(warmup, npts - 1, steps)).astype(int))
    return [int(i) for i in idx]


def run_gpu(lib, spec, precision, timed, ktimer, keep=(), ws=0, kt_kind=0):
    """Solve the sweep up to the last timed point, timing the points listed in
    `timed`; returns (seconds, currents, current floors, kernel timer,
    iteration counts of the timed points, {index: (psi, phi_n, phi_p)} of the
    converged states listed in keep, read outside the timed calls).  The
    kernel timer samples kind kt_kind (0 level-0 smoother sweep, 1 Krylov
    SpMV)."""
    from semiconductor_devices_simulator_amd.core import dptr
    cids = load_spec(lib, spec)
    assert lib.sdx_set_precision(precision) == 1
    lib.sdx_ws_mode(0)
    t0 = time.perf_counter()
    if not lib.api3_solve():
        raise RuntimeError("equilibrium solve failed")
    print(f"[bench]   precision {precision}: equilibrium {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    lib.sdx_ws_mode(ws)
    lib.sdx_ws_reset_stats()
    pts = sweep_points(spec)
    tset = set(timed)
    its = dict(pcg=0, bicgstab=0, newton=0, gummel=0)
    dt, I, F, states = 0.0, [], [], {}
    if ktimer:
        lib.sdx_ktimer_enable(kt_kind, 8)
    for i in range(0, max(timed) + 1):
        on = i in tset
        if ktimer:
            lib.sdx_ktimer_pause(0 if on else 1)       # sample the timed points only
        g0 = [lib.api3_get_stat(q) for q in range(4)]
        t0 = time.perf_counter()
        Ii = solve_points(lib, cids, pts, i, i + 1, F if on else None)[0]
        t1 = time.perf_counter()
        if on:
            dt += t1 - t0
            I.append(Ii)
            g1 = [lib.api3_get_stat(q) for q in range(4)]
            for q, key in enumerate(("pcg", "bicgstab", "newton", "gummel")):
                its[key] += int(g1[q] - g0[q])
            print(f"[bench]   point {i} ({pts[i][1]:.3g} V, hold {pts[i][2]}) {t1 - t0:.2f} s", file=sys.stderr,
                  flush=True)
        if i in keep:
            st = tuple(np.zeros(spec.N) for _ in range(3))
            lib.api3_get_state(dptr(st[0]), dptr(st[1]), dptr(st[2]))
            states[i] = st
    kt = np.zeros(4)
    if ktimer:
        lib.sdx_ktimer_read(kt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    lib.sdx_ws_mode(0)
    return dt, np.array(I), np.array(F), kt, its, states


def _floor_parity(I, Ir, Fr, drain, against):
    """|I - I_ref| <= 1e-8 |I_ref| + floor_ref on every terminal (floor_ref: the
    oracle's own resolution floor of that current, api3_get_contact_Ifloor)."""
    d = np.abs(I - Ir)
    bound = 1e-8 * np.abs(Ir) + Fr
    ok = d <= bound
    resolved = np.abs(Ir[:, drain]) > 100.0 * Fr[:, drain]
    rel = d[:, drain] / np.maximum(np.abs(Ir[:, drain]), 1e-300)
    return dict(against=against, points_within_tolerance=int(ok.all(axis=1).sum()), points=int(len(I)),
                worst_err_over_bound=float(np.max(d / np.maximum(bound, 1e-300))),
                max_rel_drain_current_err_resolved=float(rel[resolved].max()) if resolved.any() else None,
                criterion="|dI| <= 1e-8|I_ref| + floor_ref, every terminal")


def parity(I, timed, mesh, workload="mosfet"):
    """Currents of the timed points against the reference core's sweep, and
    against its sweep converged far below the default tolerances (the exact
    discrete solution to within 1e-8 + floor) when that fixture is present."""
    out = {}
    for tag, suffix, what in (("reference", "", "reference C core (FP64), default tolerances"),
                              ("reference_tight", "_tight", "reference C core (FP64), tol_p 1e-12, tol_c 1e-11")):
        path = os.path.join(ROOT, "tests", "golden", "e2e_%s_%dx%dx%d%s.npz" % (workload, *mesh, suffix))
        if os.path.isfile(path):
            g = np.load(path)
            out[tag] = _floor_parity(I, g["I"][timed], g["F"][timed], DRAIN[workload], what)
    return out or None


def mixed_vs_fp64(I32, I64, F64, st32, st64, drain=1):
    """A mixed-precision method against the plain FP64 GPU method: currents of
    every timed point and the fields of the last one."""
    out = _floor_parity(I32, I64, F64, drain, "plain FP64 GPU method")
    if st32 is not None and st64 is not None:
        out["last_point_max_abs_diff_V"] = {k: float(np.abs(a - b).max())
                                           for k, a, b in zip(("psi", "phi_n", "phi_p"), st32, st64)}
    return out


# CPU-baseline sample: at each window start the reference resumes from the
# GPU's converged state (api3_set_state), solves the next point untimed (it
# re-arms the sweep predictor), then is timed on the point after it.  The
# windows span both drain biases of the Id-Vg sweep.
CPU_WINDOWS = {96: (8, 18, 28, 38, 43), 192: (8, 28, 43)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(states, mesh, workload="mosfet"):
    ref = os.path.join(ROOT, "oracle", "_ref", "libsemisim_ref.so")
    if not os.path.isfile(ref) or not states:
        return None
    # the host cores allotted to this GPU: OMP_NUM_THREADS where the machine
    # sets it (16 per GPU on the MI355X boxes), else every core
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "states.npz")
        np.savez(path, idx=np.array(sorted(states)), mesh=np.array(mesh), workload=np.array(workload),
                 **{f"s{i}": np.stack(states[i]) for i in states})
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-worker", path],
                           env=env, capture_output=True, text=True, timeout=900)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        return None
    out = json.loads(r.stdout.strip().splitlines()[-1])
    return dict(value=round(out["points_per_s"], 4), unit="bias points/s", cores=threads, kind="reference",
                host_cpus=os.cpu_count(), cpu_model=cpu_model(),
                sample=f"{out['points']} bias points {out['timed']} (each the second point after a sweep position "
                       f"where the reference resumes from the GPU's converged state), {out['seconds']:.1f} s; "
                       f"reference C core built from its sources, OpenMP, {threads} threads",
                per_point_s=out["per_point_s"])


def cpu_baseline_worker(path):
    from semiconductor_devices_simulator_amd import core
    from semiconductor_devices_simulator_amd.core import dptr
    lib = core.bind(ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libsemisim_ref.so")))
    data = np.load(path)
    mesh = [int(v) for v in data["mesh"]]
    spec = make_spec(str(data["workload"]), mesh)
    cids = load_spec(lib, spec)
    pts = sweep_points(spec)
    dt, per, timed = 0.0, [], []
    for i in data["idx"]:
        i = int(i)
        st = data[f"s{i}"]
        tgt, v, hold = pts[i]
        for c, hv in hold.items():
            lib.api3_set_contact_V(cids[c], hv)
        lib.api3_set_contact_V(cids[tgt], v)
        psi, phin, phip = (np.ascontiguousarray(a) for a in st)
        assert lib.api3_set_state(dptr(psi), dptr(phin), dptr(phip)) == 1
        solve_points(lib, cids, pts, i + 1, i + 2)
        t0 = time.perf_counter()
        solve_points(lib, cids, pts, i + 2, i + 3)
        per.append(time.perf_counter() - t0)
        dt += per[-1]
        timed.append([round(pts[i + 2][2].get(1, 0.0), 3), round(pts[i + 2][1], 3)])
    print(json.dumps(dict(points_per_s=len(per) / dt, seconds=dt, points=len(per), per_point_s=per,
                          timed=" ".join(f"(Vd {a} V, Vg {b} V)" for a, b in timed))))


def describe_timed(pts, timed, workload):
    """The timed bias points, grouped by the held contact voltages."""
    groups = {}
    for i in timed:
        tgt, v, hold = pts[i]
        groups.setdefault(tuple(sorted(hold.items())), []).append(v)
    name = {"mosfet": ("Vd", "Vg", {1: "Vd"}), "vdmos": ("Vgs", "Vd", {2: "Vgs"}), "prism": ("Vd", "Vg", {1: "Vd"})}[workload]
    parts = []
    for hold, vs in groups.items():
        h = dict(hold)
        key = [c for c in name[2] if c in h]
        lab = f"{name[2][key[0]]} {h[key[0]]:g} V: " if key else ""
        parts.append(f"{lab}{name[1]} {min(vs):g}-{max(vs):g} V ({len(vs)} points)")
    return "; ".join(parts)


def pmc_traffic(mesh, workload="mosfet"):
    """HBM bytes per launch of the level-0 sweep from the committed counter
    summary of this mesh (two rocprofv3 --pmc runs, FETCH_SIZE and WRITE_SIZE,
    calibrated against a copy kernel), or (None, None) when there is none."""
    name = "k_line_scan_level0_%s_%dx%dx%d_pmc.md" % (workload, *mesh)
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.isfile(path):
        return None, None
    for line in open(path):
        if line.startswith("Level-0") and "MB per launch" in line:
            mb = float(line.rsplit("=", 1)[1].split("MB")[0])
            return round(mb * 1e6), "profiles/" + name
    return None, None

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mesh", type=int, nargs=3, default=None)
    ap.add_argument("--workload", choices=["mosfet", "vdmos", "prism"], default="mosfet")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--sweep-points", type=int, default=0,
                    help="solve only the first N points of the sweep (short profiling runs)")
    ap.add_argument("--cpu-baseline-worker", default="")
    ap.add_argument("--batched", action="store_true",
                    help="also time the whole sweep with its two drain-bias curves in one worker process after the "
                         "other and in two at once (sweep-level batching, both methods)")
    a = ap.parse_args()
    if a.cpu_baseline_worker:
        cpu_baseline_worker(a.cpu_baseline_worker)
        return 0
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU only (got --gpus {a.gpus}); the sweep is not sharded "
              "across GPUs", file=sys.stderr)
        return 2
    from semiconductor_devices_simulator_amd import core, devices, warmstart
    lib = core.load_core()
    a.mesh = a.mesh or DEFAULT_MESH[a.workload]
    if a.workload == "prism":
        return main_prism(a, lib)
    if a.workload == "vdmos" and a.mesh == DEFAULT_MESH["vdmos"]:
        # the blocking sweep to 44 V: this mesh breaks down at 46 V
        if a.sweep_points <= 0 or a.sweep_points > 45:
            a.sweep_points = 45
        if "--steps" not in sys.argv:
            a.steps = 43
    spec = devices.mosfet(*a.mesh) if a.workload == "mosfet" else devices.vdmos(*a.mesh)
    if a.workload != "mosfet":          # the trained network and the CPU sample are MOSFET-only
        a.no_learned = a.no_cpu_baseline = True
    npts = spec.bias_points()
    if a.sweep_points > 0:
        npts = min(npts, a.sweep_points)
    W = max(0, min(a.warmup, npts - 1))
    K = max(1, min(a.steps, npts - W))
    pts = sweep_points(spec)
    timed = timed_indices(npts, W, K)
    K = len(timed)
    workload = "%s_%dx%dx%d %s; timed %s" % (a.workload, *a.mesh, SWEEP[a.workload],
                                             describe_timed(pts, timed, a.workload))
    buf = ctypes.create_string_buffer(256)
    lib.sdx_device_name(buf, 256)
    print(f"[bench] device {buf.value.decode()}; {workload}; points {timed}", file=sys.stderr, flush=True)
    windows = [i for i in (CPU_WINDOWS.get(a.mesh[0], CPU_WINDOWS[96]) if a.workload == "mosfet" else ())
               if i + 2 < npts and i <= timed[-1]]
    last = timed[-1]
    t64, I64, F64, _, its64, states = run_gpu(lib, spec, 0, timed, ktimer=False, keep=windows + [last])
    print(f"[bench] plain FP64: {K / t64:.3f} points/s", file=sys.stderr, flush=True)
    t32, I32, _, kt, its32, st32 = run_gpu(lib, spec, 1, timed, ktimer=True, keep=[last])
    print(f"[bench] FP32 multigrid: {K / t32:.3f} points/s", file=sys.stderr, flush=True)
    learned = None
    if os.path.isfile(warmstart.MODEL) and not a.no_learned:
        meta = warmstart.load(lib, warmstart.MODEL)
        lib.sdx_ws_config(1, 0.0)       # the network's psi correction, applied at every bias point
        tL, IL, _, _, itsL, stL = run_gpu(lib, spec, 1, timed, ktimer=False, keep=[last], ws=3)
        lib.sdx_ws_config(7, 0.0)
        learned = dict(points_per_s=round(K / tL, 4), seconds=tL, iterations=itsL, warm_start=warmstart.ws_stats(lib),
                       model=dict(file=os.path.relpath(warmstart.MODEL, ROOT), samples=meta.get("samples"),
                                  heldout_loss=meta.get("heldout_loss_after")),
                       I=IL, last=stL.get(last))
        lib.sdx_nn_free()
        print(f"[bench] learned warm start + FP32 multigrid: {K / tL:.3f} points/s", file=sys.stderr, flush=True)
    roof = None
    traffic, traffic_src = pmc_traffic(a.mesh, a.workload)
    if kt[0] > 0:
        avg_s = kt[1] / kt[0] * 1e-3
        ach = kt[2] / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                    kernel="k_line_scan<float> level 0 (zebra line Gauss-Seidel sweep, one colour)",
                    bytes_per_launch=kt[2], avg_launch_us=round(avg_s * 1e6, 2), launches_timed=int(kt[0]))
    cpu = None if a.no_cpu_baseline else cpu_baseline({i: states[i] for i in windows if i in states}, a.mesh)
    # the metric's method: the learned warm start on the FP32 multigrid when it
    # kept corrections and beat the FP32 multigrid alone; otherwise the FP32
    # multigrid (the learned run is reported beside it either way)
    ws_kept = learned["warm_start"].get("kept", 0) if learned else 0
    use_learned = bool(learned) and ws_kept > 0 and learned["seconds"] < t32
    t_val = learned["seconds"] if use_learned else t32
    value = K / t_val
    D = DRAIN[a.workload]
    out = {
        "metric": "bias points/s at 1e-8 rel current err, learned+FP32-AMG vs plain FP64 on GPU",
        "value": round(value, 4),
        "unit": "bias points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1e3 * t_val / K, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (residuals, dots, Newton, assembly, Bernoulli) + f32 (multigrid, inner Krylov vectors)",
        "data": "synthetic",
        "config": {"workload": workload, "nodes": spec.N, "mesh": list(spec.shape),
                   "timed_points": [[round(pts[i][2].get(D if a.workload == "mosfet" else 2, 0.0), 4),
                                     round(pts[i][1], 4)] for i in timed],
                   "timed_point_fields": ["Vd", "Vg"] if a.workload == "mosfet" else ["Vgs", "Vd"],
                   "method": "learned warm start + FP32-AMG" if use_learned else "FP32-AMG",
                   "learned_warm_start": use_learned},
        "plain_fp64_points_per_s": round(K / t64, 4),
        "fp32_mg_points_per_s": round(K / t32, 4),
        "learned_fp32_mg_points_per_s": learned["points_per_s"] if learned else None,
        "speedup_vs_plain_fp64": round(t64 / t_val, 3),
        "iterations": {"fp32_mg": its32, "fp64": its64, "learned_fp32_mg": learned["iterations"] if learned else None},
        "warm_start": ({k: learned[k] for k in ("warm_start", "model")} if learned else None),
        "parity": {"fp32_mg": parity(I32, timed, a.mesh, a.workload), "fp64": parity(I64, timed, a.mesh, a.workload),
                   "learned_fp32_mg": parity(learned["I"], timed, a.mesh, a.workload) if learned else None,
                   "fp32_mg_vs_fp64": mixed_vs_fp64(I32, I64, F64, st32.get(last), states.get(last), D),
                   "learned_fp32_mg_vs_fp64": (mixed_vs_fp64(learned["I"], I64, F64, learned["last"], states.get(last),
                                                             D) if learned else None)},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if a.batched and a.workload == "mosfet":
        out["batched"] = batched(a.mesh, spec)
    print(json.dumps(out))
    return 0


def batched(mesh, spec):
    """Sweep-level batching: the whole Id-Vg sweep (both drain-bias curves,
    every point) in one worker process (the curves one after the other) and
    in two worker processes at once, each curve from the equilibrium
    solution, in both methods; currents of every point against the
    reference's sweeps."""
    from semiconductor_devices_simulator_amd import batch
    n = spec.bias_points()
    res = {}
    for prec, tag in ((1, "fp32_mg"), (0, "fp64")):
        res[tag] = {}
        for label, groups in (("one_worker", [[0, 1]]), ("two_workers", [[0], [1]])):
            o = batch.run_concurrent(("mosfet", tuple(mesh), {}), groups, precision=prec)
            R = batch.concat(o["results"])
            res[tag][label] = dict(points=o["points"], seconds=round(o["seconds"], 3),
                                   points_per_s=round(o["points_per_s"], 4),
                                   gummel=int(sum(int(r["iterations"][3]) for r in o["results"])),
                                   converged=int(np.sum(R["conv"])),
                                   parity=parity(R["I"], list(range(n)), mesh) if o["points"] == n else None)
            print(f"[bench] batched {tag} {label}: {o['points_per_s']:.3f} points/s", file=sys.stderr, flush=True)
        res[tag]["gain"] = round(res[tag]["two_workers"]["points_per_s"] / res[tag]["one_worker"]["points_per_s"], 3)
    return dict(workers=2, groups=["Vd 0.05 V curve", "Vd 1.0 V curve"], **res)


def main_prism(a, lib):
    """Supplementary line: the graph (prism-mesh) path, FP64 aggregation
    multigrid, with the graph Krylov SpMV as the roofline kernel and the
    reference core's graph path on the host cores beside it."""
    spec = make_spec("prism", a.mesh)
    pts = sweep_points(spec)
    npts = len(pts)
    W = max(0, min(a.warmup, npts - 1))
    K = max(1, min(a.steps, npts - W))
    timed = timed_indices(npts, W, K)
    K = len(timed)
    workload = "prism_mosfet_%dx%dx%d %s; %d nodes, %d directed edges; timed %s" % (
        *a.mesh, SWEEP["prism"], spec.N, spec.E, describe_timed(pts, timed, "prism"))
    print(f"[bench] {workload}; points {timed}", file=sys.stderr, flush=True)
    windows = [i for i in (3, 8) if i + 2 < npts and i <= timed[-1]]
    t, I, F, kt, its, states = run_gpu(lib, spec, 0, timed, ktimer=True, keep=windows, kt_kind=1)
    print(f"[bench] graph path: {K / t:.3f} points/s", file=sys.stderr, flush=True)
    roof = None
    if kt[0] > 0:
        avg_s = kt[1] / kt[0] * 1e-3
        ach = kt[2] / avg_s / 1e9
        roof = dict(bound="hbm", achieved=round(ach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(ach / HBM_PEAK_GBS, 4), traffic=None,
                    kernel="k_gspmv_dots (graph CSR Krylov product fused with its dots)",
                    bytes_per_launch=kt[2], avg_launch_us=round(avg_s * 1e6, 2), launches_timed=int(kt[0]))
    cpu = None if a.no_cpu_baseline else cpu_baseline({i: states[i] for i in windows if i in states}, a.mesh,
                                                      "prism")
    out = {
        "metric": "bias points/s at 1e-8 rel current err, graph (prism-mesh) path on GPU",
        "value": round(K / t, 4), "unit": "bias points/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(1e3 * t / K, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "nodes": spec.N, "edges": spec.E, "mesh": list(a.mesh),
                   "timed_points": [[round(pts[i][2].get(1, 0.0), 4), round(pts[i][1], 4)] for i in timed],
                   "timed_point_fields": ["Vd", "Vg"], "method": "FP64 aggregation multigrid (graph)"},
        "iterations": its,
        "parity": parity(I, timed, a.mesh, "prism_mosfet"),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if a.batched and a.workload == "mosfet":
        out["batched"] = batched(a.mesh, spec)
    print(json.dumps(out))
    return 0


def batched(mesh, spec):
    """Sweep-level batching: the whole Id-Vg sweep (both drain-bias curves,
    every point) in one worker process (the curves one after the other) and
    in two worker processes at once, each curve from the equilibrium
    solution, in both methods; currents of every point against the
    reference's sweeps."""
    from semiconductor_devices_simulator_amd import batch
    n = spec.bias_points()
    res = {}
    for prec, tag in ((1, "fp32_mg"), (0, "fp64")):
        res[tag] = {}
        for label, groups in (("one_worker", [[0, 1]]), ("two_workers", [[0], [1]])):
            o = batch.run_concurrent(("mosfet", tuple(mesh), {}), groups, precision=prec)
            R = batch.concat(o["results"])
            res[tag][label] = dict(points=o["points"], seconds=round(o["seconds"], 3),
                                   points_per_s=round(o["points_per_s"], 4),
                                   gummel=int(sum(int(r["iterations"][3]) for r in o["results"])),
                                   converged=int(np.sum(R["conv"])),
                                   parity=parity(R["I"], list(range(n)), mesh) if o["points"] == n else None)
            print(f"[bench] batched {tag} {label}: {o['points_per_s']:.3f} points/s", file=sys.stderr, flush=True)
        res[tag]["gain"] = round(res[tag]["two_workers"]["points_per_s"] / res[tag]["one_worker"]["points_per_s"], 3)
    return dict(workers=2, groups=["Vd 0.05 V curve", "Vd 1.0 V curve"], **res)


if __name__ == "__main__":
    sys.exit(main())
