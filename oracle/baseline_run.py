"""CPU baseline of the 4K workload: the oracle's assembly + Jacobi-PCG to
1e-6 on the host, built for the host it runs on (oracle/Makefile target
`baseline`: gcc -O3 -march=native) and timed as the median of full solves.

    OMP_PROC_BIND=close OMP_NUM_THREADS=16 python -m oracle.baseline_run [--runs 3]

bench.py starts this in a child process so that the OpenMP settings apply
before any OpenMP runtime starts.  Prints one JSON line.  Reported beside the
GPU result; it is not the GPU path.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from gpumatting_amd import synth  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def cpu_info():
    model, flags = "", ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and not model:
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return model, flags


def build_for_host():
    model, flags = cpu_info()
    tag = hashlib.sha1((model + flags).encode()).hexdigest()[:12]
    out = os.path.join("_native", f"libbaseline-{tag}.so")
    subprocess.run(["make", "-s", "-C", HERE, "baseline", f"OUT={out}"], check=True)
    return os.path.join(HERE, out), model


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--tol", type=float, default=1e-6)
    args = ap.parse_args()
    path, model = build_for_host()
    L = O.baseline_lib(path)
    p = synth.config_4k()
    W, H, N = p.W, p.H, p.N
    bands = np.ascontiguousarray(O.bands25(W), np.int32)
    pad = O.pad_for(W)
    im = np.ascontiguousarray(p.image, np.float32)
    lam = np.ascontiguousarray(p.lam, np.float32).ravel()
    tgt = np.ascontiguousarray(p.target, np.float32).ravel()
    # every buffer allocated and touched before timing
    a = np.ones((25, N), np.float32)
    b = np.ones(N, np.float32)
    dinv = np.ones(N, np.float32)
    xp = np.zeros(N + 2 * pad, np.float32)
    runs = []
    for _ in range(args.runs):
        xp[:] = 0
        t0 = time.perf_counter()
        L.orc_levin_gather_acc_f32(a, N, b, dinv.ctypes.data, im, 3, W, lam.ctypes.data, tgt.ctypes.data,
                                   C.c_float(p.eps), W, H)
        t1 = time.perf_counter()
        it = C.c_int(0)
        rr = L.orc_baseline_pcg_f32(xp.ctypes.data + 4 * pad, a.ctypes.data, N, bands, 25, N, b.ctypes.data, pad,
                                    100000, args.tol, C.byref(it))
        t2 = time.perf_counter()
        runs.append((t2 - t0, t1 - t0, t2 - t1, it.value, rr))
    med = statistics.median(r[0] for r in runs)
    m = min(runs, key=lambda r: abs(r[0] - med))
    print(json.dumps({"ms": 1e3 * med, "asm_ms": 1e3 * m[1], "pcg_ms": 1e3 * m[2], "iterations": m[3],
                      "relres": m[4], "runs": [round(1e3 * r[0], 1) for r in runs], "threads": O.num_threads(),
                      "proc_bind": os.environ.get("OMP_PROC_BIND", ""), "cpu": model}))


if __name__ == "__main__":
    main()
