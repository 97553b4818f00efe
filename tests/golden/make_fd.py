"""Writes tests/golden/fd_grad_512.npz: FP64 central finite differences
(h = 1e-4) of loss = ||alpha - alpha_gt||^2 at 256 random coordinates of the
first image of the differentiation workload (512x512, synth.config_grad()[0],
eps 1e-7): 192 RGB channels (8 of them on the image border) and 64 per-pixel
constraint weights (32 constrained pixels, 32 unconstrained).  Computed by the
oracle (oracle.pyoracle.matte_loss_fd_f64); about five minutes on 8 cores.

    python tests/golden/make_fd.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gpumatting_amd import synth  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

H_STEP = 1e-4


def coordinates(p, seed=17):
    rng = np.random.default_rng(seed)
    W, H = p.W, p.H
    rgb = []
    border = [(0, 0), (0, W - 1), (H - 1, 0), (H - 1, W - 1), (0, W // 2), (H - 1, W // 3), (H // 2, 0),
              (H // 3, W - 1)]
    for y, x in border:
        rgb.append((y, x, int(rng.integers(0, 3))))
    while len(rgb) < 192:
        rgb.append((int(rng.integers(0, H)), int(rng.integers(0, W)), int(rng.integers(0, 3))))
    lam = p.lam.ravel()
    con = rng.choice(np.flatnonzero(lam > 0), 32, replace=False)
    unc = rng.choice(np.flatnonzero(lam == 0), 32, replace=False)
    return np.array(rgb, np.int64), np.concatenate([con, unc]).astype(np.int64)


def main():
    p = synth.config_grad(1)[0]
    rgb, lam_idx = coordinates(p)
    coords = [("rgb", int(y), int(x), int(c)) for y, x, c in rgb] + [("lam", int(i)) for i in lam_idx]
    t = time.time()
    fd = O.matte_loss_fd_f64(p.image, p.lam, p.target, p.alpha_gt, coords, h=H_STEP, eps=1e-7)
    print(f"{len(coords)} derivatives in {time.time() - t:.0f} s")
    np.savez_compressed(os.path.join(HERE, "fd_grad_512.npz"), rgb=rgb, lam_idx=lam_idx, fd_rgb=fd[:len(rgb)],
                        fd_lam=fd[len(rgb):], h=np.float64(H_STEP), W=p.W, H=p.H, seed=5000)


if __name__ == "__main__":
    main()
