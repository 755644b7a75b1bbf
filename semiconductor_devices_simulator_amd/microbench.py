"""Kernel microbenchmark of the HIP core on a solved device.

Solves a workload device at one bias point, then times back-to-back launches
of one kernel on the solver stream (HIP events, sdx_microbench) and prints
ms/launch with its algorithmic bytes and GB/s.  Used for the roofline and
rocprofv3 counter runs:

    python -m semiconductor_devices_simulator_amd.microbench --mesh 96 64 64 \\
        --precision 1 --kernel line_y --iters 200

kernels: line_x line_y line_z (one colour of a level-0 zebra line
Gauss-Seidel sweep), rbgs (one colour of a point red-black sweep), spmv
(Krylov SpMV fused with a dot), restrict (residual + restriction level 0 to
1), vcycle, copy (N doubles, a bandwidth reference), poisson (Poisson
assembly), cont (continuity assembly), copy_big (8M doubles, 64 MB each way:
larger than the L2 of all XCDs together, the calibration reference of the
rocprofv3 FETCH_SIZE / WRITE_SIZE counters), precond (the preconditioner as
the FP64 Krylov solvers call it: precision conversions and the V-cycle), roundtrip (a dot product whose value the host
reads back, host wall clock per round trip).

With --sweep it instead times the workload's whole Id-Vg sweep in one
method (no kernel launches of its own), the run whose kernel trace
(rocprofv3 --kernel-trace) tracestat.py summarises: the kernels' busy time
against this wall time says how much of the sweep the GPU idles.

    python -m semiconductor_devices_simulator_amd.microbench --mesh 192 128 128 --precision 1 --sweep
"""
import argparse
import os
import json
import sys

from . import core, devices

KERNELS = ["line_x", "line_y", "line_z", "rbgs", "spmv", "restrict", "vcycle", "copy", "poisson", "cont", "copy_big",
           "precond", "roundtrip", "line_x_l1", "line_y_l1", "line_z_l1", "spmv_staged"]


def algorithmic_bytes(kernel, N, precision):
    w = 4 if precision == 1 else 8
    # level-0 line sweep: b, the other colour's x and the x written in T, 4
    # couplings and 3 line factors in T, or in BF16 in the FP32 method
    lw = 3 * w + 7 * (2 if precision == 1 and os.environ.get("SEMISIM_MG_BF16", "1") != "0" else w)
    return {
        "line_x": N / 2 * lw, "line_y": N / 2 * lw, "line_z": N / 2 * lw,
        "rbgs": N / 2 * 10 * w, "spmv": N * 10 * 8, "restrict": N / 2 * 9 * w + N / 8 * w,
        "vcycle": None, "copy": N * 16, "poisson": N * 161, "cont": N * 190, "copy_big": (8 << 20) * 16, "precond": None, "roundtrip": None,
        "line_x_l1": N / 16 * 10 * w, "line_y_l1": N / 16 * 10 * w, "line_z_l1": N / 16 * 10 * w,
        "spmv_staged": N * (10 * 8 + w),
    }[kernel]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--mesh", type=int, nargs=3, default=[96, 64, 64])
    ap.add_argument("--precision", type=int, default=1)
    ap.add_argument("--kernel", default="line_y", choices=KERNELS + ["all"])
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--sweep", action="store_true", help="time the whole Id-Vg sweep in one method")
    ap.add_argument("--tune", action="append", default=[],
                    help="WHAT=VALUE passed to sdx_tune before the solve (4=0: the k_line_gs line kernel)")
    a = ap.parse_args(argv)
    lib = core.load_core()
    spec = devices.mosfet(*a.mesh)
    cids = devices.load_device(lib, spec)
    for t in a.tune:              # after api3_init, which reads the environment knobs
        w, val = (int(z) for z in t.split("="))
        lib.sdx_tune(w, val)
    lib.sdx_set_precision(a.precision)
    if not lib.api3_solve():
        raise RuntimeError("equilibrium solve failed")
    if a.sweep:
        return time_sweep(lib, spec, cids, a)
    for c, v in ((1, 1.0), (2, 0.6)):
        lib.api3_set_contact_V(cids[c], v)
    if not lib.api3_resolve():
        raise RuntimeError("bias point solve failed")
    names = KERNELS if a.kernel == "all" else [a.kernel]
    for name in names:
        ms = lib.sdx_microbench(KERNELS.index(name), a.iters)
        by = algorithmic_bytes(name, spec.N, a.precision)
        out = dict(kernel=name, mesh=a.mesh, precision=a.precision, us_per_launch=round(ms * 1e3, 3),
                   bytes_per_launch=by, GBps=None if by is None else round(by / ms / 1e6, 1))
        print(json.dumps(out), flush=True)
    return 0


def time_sweep(lib, spec, cids, a):
    """The whole sweep of the workload device, from its equilibrium."""
    import time
    import numpy as np
    from . import sweep
    g0 = [lib.api3_get_stat(q) for q in range(4)]
    t0 = time.perf_counter()
    R = sweep.run_sweep(lib, spec, cids)
    dt = time.perf_counter() - t0
    its = [int(lib.api3_get_stat(q) - g) for q, g in enumerate(g0)]
    out = dict(sweep=spec.name, mesh=a.mesh, precision=a.precision, points=len(R["V"]),
               converged=int(np.sum(R["conv"])), seconds=round(dt, 3), points_per_s=round(len(R["V"]) / dt, 4),
               iterations=dict(zip(("pcg", "bicgstab", "newton", "gummel"), its)))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
