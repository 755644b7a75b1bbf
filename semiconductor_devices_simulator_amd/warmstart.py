"""Learned warm start: training data from the simulator's own sweeps, training
on the GPU, and use in sweeps.

The network (csrc/nn.cpp, a 3-D convolutional U-Net on a 64^3 grid with BF16
matrix-core kernels) reads the plain continuation's predicted state of a bias
point - psi, phi_n, phi_p resampled from the mesh in index space, the signed
doping log-magnitude and the bias step - and predicts the correction to the
converged state.  In use (sdx_ws_mode(2)) the engine applies the correction
and keeps it only if it lowers the starting Poisson residual; Newton/Gummel
then converge to the usual tolerances either way.

Training data: Id-Vg sweeps of random MOSFETs (devices.random_mosfet), each
bias point reached in one continuation step giving one (prediction, converged
state) pair, recorded on the device (sdx_ws_mode(1)).

    python -m semiconductor_devices_simulator_amd.warmstart train --devices 48 --steps 3000 --out model.npz
    python -m semiconductor_devices_simulator_amd.warmstart eval --model model.npz --seeds 10000 10001
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

from . import core, devices, sweep
from .core import fptr

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "warmstart_unet64.npz")
GRID = 64
CHANNELS = (32, 32, 64, 64)
SCALE = 1e-3          # V per unit of network output


def init_network(lib, grid=GRID, channels=CHANNELS, seed=0, scale=SCALE):
    if lib.sdx_nn_init(grid, *channels, seed) != 1:
        raise RuntimeError("network allocation failed")
    lib.sdx_nn_set_scale(scale)


def save(lib, path, meta=None):
    n = lib.sdx_nn_nparams()
    p = np.zeros(n, dtype=np.float32)
    if lib.sdx_nn_get_params(fptr(p), n) != 1:
        raise RuntimeError("reading the parameters failed")
    np.savez(path, params=p, grid=GRID, channels=np.array(CHANNELS), scale=lib.sdx_nn_get_scale(),
             meta=json.dumps(meta or {}))


def load(lib, path=MODEL):
    """Create the network and load trained parameters (no pickles involved)."""
    z = np.load(path, allow_pickle=False)
    init_network(lib, int(z["grid"]), tuple(int(c) for c in z["channels"]), 0, float(z["scale"]))
    p = np.ascontiguousarray(z["params"], dtype=np.float32)
    if lib.sdx_nn_set_params(fptr(p), len(p)) != 1:
        raise RuntimeError(f"{path}: parameter count does not match the network")
    return json.loads(str(z["meta"])) if "meta" in z else {}


def ws_stats(lib):
    out = np.zeros(4)
    lib.sdx_ws_stats(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return dict(tried=int(out[0]), kept=int(out[1]), rejected=int(out[2]), recorded=int(out[3]))


def capture(lib, seeds, mesh, log=print):
    """Record training samples from the sweeps of random MOSFETs."""
    lib.sdx_ws_mode(1)
    for q, seed in enumerate(seeds):
        spec = devices.random_mosfet(seed, *mesh)
        cids = devices.load_device(lib, spec)
        t0 = time.perf_counter()
        if lib.api3_solve() != 1:
            log(f"[ws] seed {seed}: equilibrium failed, skipped")
            continue
        before = lib.sdx_nn_dataset_size()
        R = sweep.run_sweep(lib, spec, cids)
        log(f"[ws] device {q + 1}/{len(seeds)} seed {seed}: {int(np.sum(R['conv']))}/{len(R['conv'])} points, "
            f"{lib.sdx_nn_dataset_size() - before} samples, {time.perf_counter() - t0:.1f} s")
    lib.sdx_ws_mode(0)
    return lib.sdx_nn_dataset_size()


def train(lib, steps, n_train, lr=1e-3, seed=1, log=print, chunk=250):
    losses = []
    for s0 in range(0, steps, chunk):
        k = min(chunk, steps - s0)
        buf = np.zeros(k, dtype=np.float32)
        lr_k = lr * (0.1 if s0 >= 0.8 * steps else 1.0)
        m = lib.sdx_nn_train(k, n_train, lr_k, seed + s0, fptr(buf))
        if m < 0:
            raise RuntimeError("training step failed")
        losses.append(buf)
        log(f"[ws] steps {s0 + k}/{steps}: mean loss {m:.4e}")
    return np.concatenate(losses)


def sweep_with(lib, spec, mode, precision=1):
    """Solve the spec's sweep with the warm start off (0) or on (2); returns
    currents, Gummel/Newton/Krylov iterations, wall time and warm-start stats."""
    cids = devices.load_device(lib, spec)
    lib.sdx_set_precision(precision)
    lib.sdx_ws_mode(0)
    if lib.api3_solve() != 1:
        raise RuntimeError("equilibrium failed")
    lib.sdx_ws_mode(mode)
    lib.sdx_ws_reset_stats()
    g0 = [lib.api3_get_stat(i) for i in range(4)]
    t0 = time.perf_counter()
    R = sweep.run_sweep(lib, spec, cids)
    dt = time.perf_counter() - t0
    g1 = [lib.api3_get_stat(i) for i in range(4)]
    lib.sdx_ws_mode(0)
    return dict(I=R["I"], F=R["F"], conv=R["conv"], seconds=dt,
                its=dict(zip(("pcg", "bicgstab", "newton", "gummel"), [b - a for a, b in zip(g0, g1)])),
                ws=ws_stats(lib))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--devices", type=int, default=48)
    t.add_argument("--first-seed", type=int, default=0)
    t.add_argument("--mesh", type=int, nargs=3, default=[96, 64, 64])
    t.add_argument("--steps", type=int, default=3000)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--holdout", type=float, default=0.1)
    t.add_argument("--out", default=MODEL)
    e = sub.add_parser("eval")
    e.add_argument("--model", default=MODEL)
    e.add_argument("--seeds", type=int, nargs="*", default=[10000, 10001])
    e.add_argument("--mesh", type=int, nargs=3, default=[96, 64, 64])
    e.add_argument("--bench-mesh", type=int, nargs=3, default=None)
    e.add_argument("--modes", type=int, nargs="*", default=[2],
                   help="2: keep a correction only if it lowers the starting residual, 3: always keep it")
    e.add_argument("--config", nargs="*", default=["7,0"],
                   help="FIELDS,THR pairs: fields corrected (psi 1, phi_n 2, phi_p 4), smallest correction in V")
    e.add_argument("--accept", type=float, nargs="*", default=[0.5],
                   help="mode 2: keep a correction whose starting Poisson residual is below this x the plain one's")
    a = ap.parse_args(argv)
    lib = core.load_core()
    log = lambda m: print(m, flush=True)
    if a.cmd == "train":
        init_network(lib)
        per_dev = 2 * 25
        if lib.sdx_nn_dataset_alloc(a.devices * per_dev) != 1:
            return 1
        seeds = list(range(a.first_seed, a.first_seed + a.devices))
        n = capture(lib, seeds, a.mesh, log)
        n_train = max(1, int(round(n * (1.0 - a.holdout))))
        log(f"[ws] {n} samples ({n_train} training, {n - n_train} held out)")
        v0 = lib.sdx_nn_eval(n_train, n)
        L = train(lib, a.steps, n_train, a.lr, log=log)
        v1 = lib.sdx_nn_eval(n_train, n)
        t0 = lib.sdx_nn_eval(0, min(n_train, 200))
        meta = dict(samples=n, n_train=n_train, seeds=[seeds[0], seeds[-1]], mesh=a.mesh, steps=a.steps,
                    lr=a.lr, heldout_loss_before=v0, heldout_loss_after=v1, train_loss_after=t0,
                    train_loss_first100=float(L[:100].mean()), train_loss_last100=float(L[-100:].mean()))
        save(lib, a.out, meta)
        log("[ws] " + json.dumps(meta))
        return 0
    meta = load(lib, a.model)
    log(f"[ws] model {a.model}: {json.dumps(meta)}")
    specs = [devices.random_mosfet(s, *a.mesh) for s in a.seeds]
    if a.bench_mesh:
        specs.append(devices.mosfet(*a.bench_mesh))
    variants = [tuple(float(v) for v in c.split(",")) for c in a.config]
    for spec in specs:
        r0 = sweep_with(lib, spec, 0)
        for mode, acc in [(m, x) for m in a.modes for x in (a.accept if m == 2 else [0.5])]:
            lib.sdx_ws_accept(acc)
            for chm, thr in variants:
                lib.sdx_ws_config(int(chm), thr)
                r1 = sweep_with(lib, spec, mode)
                d = np.abs(r1["I"] - r0["I"])
                ok = bool(np.all(d <= 1e-8 * np.abs(r0["I"]) + r0["F"] + r1["F"]))
                # relative deviation where the current stands well above its floors
                big = np.abs(r0["I"]) > 100.0 * (r0["F"] + r1["F"])
                dev = float(np.max(np.where(big, d / np.where(big, np.abs(r0["I"]), 1.0), 0.0)))
                log("[ws] " + json.dumps(dict(device=spec.name, mode=mode, accept=acc, fields=int(chm), thr=thr,
                                              plain=r0["its"],
                                              learned=r1["its"], ws=r1["ws"],
                                              seconds=[round(r0["seconds"], 2), round(r1["seconds"], 2)],
                                              currents_agree=ok, max_rel_current_dev=dev)))
    lib.sdx_ws_config(7, 0.0)
    lib.sdx_ws_accept(0.5)
    return 0


if __name__ == "__main__":
    sys.exit(main())
