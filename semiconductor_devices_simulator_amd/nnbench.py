"""Kernel time of the warm-start network's training and inference (no
device solve): the U-Net at its configured size on synthetic samples, a
number of Adam steps on the device (forward, backward, weight gradients) and
of inference passes, timed on the host; the run a rocprofv3 kernel trace
profiles for the convolution kernels' matrix-core throughput.  `flops` lists
the multiply-adds of every convolution of one forward pass (2 per MAC, the
padded channel counts the kernels compute).

    python -m semiconductor_devices_simulator_amd.nnbench --steps 20 --infer 20
"""
import argparse
import ctypes
import json
import sys
import time

import numpy as np

from . import core, warmstart


def conv_flops(G=warmstart.GRID, ch=warmstart.CHANNELS, cin=16, cout=32):
    """(grid, K, N, flops) of the eight convolutions of one forward pass."""
    c0, c1, c2, c3 = ch
    spec = [(G, cin, c0), (G // 2, c0, c1), (G // 4, c1, c2), (G // 8, c2, c3), (G // 4, c2, c1),
            (G // 2, c1, c0), (G, c0, c0), (G, c0, cout)]
    return [(g, k, n, 2.0 * 27 * g ** 3 * k * n) for g, k, n in spec]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--infer", type=int, default=20)
    ap.add_argument("--samples", type=int, default=4)
    a = ap.parse_args(argv)
    lib = core.load_core()
    warmstart.init_network(lib)
    G = warmstart.GRID
    V = G ** 3
    rng = np.random.default_rng(0)
    f = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.sdx_nn_dataset_alloc(a.samples) == 1
    for _ in range(a.samples):
        x = rng.standard_normal(8 * V).astype(np.float32)
        y = rng.standard_normal(3 * V).astype(np.float32)
        assert lib.sdx_nn_dataset_put(f(x), f(y)) >= 0
    loss = np.zeros(max(1, a.steps), np.float32)
    lib.sdx_nn_train(1, a.samples, 1e-4, 1, f(loss))            # warm-up
    t0 = time.perf_counter()
    lib.sdx_nn_train(a.steps, a.samples, 1e-4, 2, f(loss))
    t_train = time.perf_counter() - t0
    x = rng.standard_normal(8 * V).astype(np.float32)
    out = np.zeros(3 * V, np.float32)
    lib.sdx_nn_forward_host(f(x), f(out))
    t0 = time.perf_counter()
    for _ in range(a.infer):
        lib.sdx_nn_forward_host(f(x), f(out))
    t_inf = time.perf_counter() - t0
    fl = conv_flops()
    print(json.dumps(dict(grid=G, channels=list(warmstart.CHANNELS), train_steps=a.steps,
                          ms_per_train_step=round(1e3 * t_train / max(1, a.steps), 3), infer=a.infer,
                          ms_per_inference=round(1e3 * t_inf / max(1, a.infer), 3),
                          forward_gflop=round(sum(x[3] for x in fl) / 1e9, 2),
                          flops=[dict(grid=g, K=k, N=n, gflop=round(v / 1e9, 3)) for g, k, n, v in fl])))
    lib.sdx_nn_free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
