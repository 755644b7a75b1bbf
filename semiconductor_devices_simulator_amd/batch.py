"""Sweep-level batching: independent bias sweeps of one device run at the
same time on one GPU, one worker process per sweep.

The api3 ABI is a singleton per process, as the reference's is
(semiconductor_3d.c keeps one device; api3_iv_sweep2 :4958-4978 runs one
sweep), so a batch is a set of processes, each with its own core instance,
device copy and HIP queues.  The GPU runs their kernels concurrently: while
one sweep is in its latency-bound phases (coarse multigrid levels, one-
workgroup reductions, host round trips) the other's bandwidth-bound level-0
kernels fill the chip.  The natural units are the curves of a family, for
example the two drain biases of an Id-Vg sweep: they share nothing, and each
starts from the equilibrium solution, as the reference's experiment engine
starts every curve (main.py:3638-3734).

    from semiconductor_devices_simulator_amd import batch
    out = batch.run_concurrent(("mosfet", (96, 64, 64), {}), [[0], [1]], precision=1)

Workers are started with the "spawn" method from a parent that need not
touch the GPU; each solves its equilibrium, waits at a common barrier, then
runs its segments.  The wall time is taken from the barrier to the last
worker's end.
"""
import multiprocessing as mp
import os
import queue
import sys
import time
import traceback

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sub_segment(spec, seg, lo=0, hi=None):
    """Points lo..hi-1 of segment seg of spec.sweep as a segment of its own
    (the same voltages)."""
    tgt, vs, ve, npts, hold = spec.sweep[seg]
    V = np.linspace(vs, ve, npts)[lo:hi]
    return (tgt, float(V[0]), float(V[-1]), len(V), hold)


def _worker(device, segs, precision, setup, barrier, out, wid):
    try:
        sys.path.insert(0, _ROOT)
        from semiconductor_devices_simulator_amd import core, devices, sweep
        name, args, kw = device
        spec = getattr(devices, name)(*args, **kw)
        segments = [sub_segment(spec, *s) if isinstance(s, tuple) else spec.sweep[s] for s in segs]
        lib = core.load_core()
        cids = devices.load_device(lib, spec)
        for what, value in setup:
            lib.sdx_tune(what, value)
        if lib.sdx_set_precision(precision) != 1:
            raise RuntimeError("precision %d rejected" % precision)
        if lib.api3_solve() != 1:
            raise RuntimeError("equilibrium solve failed")
        barrier.wait(timeout=600)
        g0 = [lib.api3_get_stat(q) for q in range(4)]
        t0 = time.perf_counter()
        R = sweep.run_sweep(lib, spec, cids, segments=segments)
        t1 = time.perf_counter()
        R["iterations"] = np.asarray([lib.api3_get_stat(q) - g for q, g in enumerate(g0)])
        out.put((wid, "ok", t0, t1, R))
    except Exception:
        out.put((wid, "error", 0.0, 0.0, traceback.format_exc()))
        try:
            barrier.abort()
        except Exception:
            pass


def run_concurrent(device, groups, precision=1, setup=(), timeout=1800):
    """Run the sweep groups of one device at the same time, one worker
    process each.  device = (builder name in devices, args, kwargs); groups =
    list of segment lists, each entry an index into spec.sweep or a tuple
    (segment, lo, hi) for points lo..hi-1 of it; setup = (what, value) pairs
    passed to sdx_tune in every worker.  Returns dict: results (per group,
    sweep.run_sweep's arrays plus the four api3_get_stat iteration counts),
    seconds (barrier to the last worker's end), points, points_per_s."""
    ctx = mp.get_context("spawn")
    barrier = ctx.Barrier(len(groups))
    out = ctx.Queue()
    procs = [ctx.Process(target=_worker, args=(device, g, precision, tuple(setup), barrier, out, w), daemon=True)
             for w, g in enumerate(groups)]
    for p in procs:
        p.start()
    res = {}
    try:
        deadline = time.monotonic() + timeout
        while len(res) < len(procs):
            try:
                wid, st, t0, t1, R = out.get(timeout=max(1.0, deadline - time.monotonic()))
            except queue.Empty:
                raise RuntimeError("batch workers did not finish within %d s" % timeout)
            if st != "ok":
                raise RuntimeError("batch worker %d failed:\n%s" % (wid, R))
            res[wid] = (t0, t1, R)
    finally:
        for p in procs:
            p.join(timeout=30)
            if p.is_alive():
                p.terminate()
                p.join(timeout=10)
    t0 = min(r[0] for r in res.values())
    t1 = max(r[1] for r in res.values())
    results = [res[w][2] for w in range(len(procs))]
    pts = sum(len(R["V"]) for R in results)
    return dict(results=results, seconds=t1 - t0, points=pts, points_per_s=pts / (t1 - t0))


def concat(results):
    """The groups' sweeps as one run (points in group order)."""
    keys = ("V", "I", "F", "conv", "t", "gummel")
    return {k: np.concatenate([np.asarray(R[k]) for R in results]) for k in keys}
