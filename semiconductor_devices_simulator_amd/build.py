"""Build the HIP solver core libsemisim_hip.so for gfx950 (in-tree).

    python -m semiconductor_devices_simulator_amd.build

Compiles csrc/*.hip and csrc/*.cpp with hipcc into lib/libsemisim_hip.so.
Every HIP translation unit is built with -Xclang -target-feature -Xclang
-packed-fp32-ops disabled (no packed FP32 instructions are emitted).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
OUT = os.path.join(LIB_DIR, "libsemisim_hip.so")
ARCH = os.environ.get("SEMISIM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value"]
SOURCES = ["k_linalg.hip", "k_krylov32.hip", "k_physics.hip", "k_nn.hip", "k_csr.hip", "k_graph.hip", "engine.cpp", "graph.cpp", "nn.cpp",
           "api.cpp"]


def build(verbose=False, jobs=4):
    os.makedirs(LIB_DIR, exist_ok=True)
    objs, procs = [], []
    try:
        for src in SOURCES:
            obj = os.path.join(LIB_DIR, os.path.splitext(src)[0] + ".o")
            lang = [] if src.endswith(".hip") else ["-x", "hip"]
            cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
            objs.append(obj)
            if len(procs) >= jobs:
                _wait(procs.pop(0))
        while procs:
            _wait(procs.pop(0))
        tmp = OUT + ".tmp"
        link = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
        _run(link, "link")
        os.replace(tmp, OUT)
    finally:
        # a failed compile must not leave the other compilers running or
        # their partial outputs behind
        for _, p in procs:
            p.communicate()
        for f in os.listdir(LIB_DIR):
            if f.endswith(".o") or ".o.tmp" in f:
                os.remove(os.path.join(LIB_DIR, f))
    return OUT


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = out.decode(errors="replace")
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{text}")


def _run(cmd, what):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stdout.decode(errors='replace')}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
