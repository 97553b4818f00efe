"""Generate tests/golden/data05_crop.npz and data05_full.npz from the
reference's sample data.

Reads /root/reference/data/05.ppm, 05_scribs.pgm and 05_gt.pgm (the input of
the reference's only end-to-end run, README.md:27), keeps a 96x64 crop that
holds foreground scribbles, background scribbles and unknown pixels, and
stores the raw 8-bit crop plus the digest of the oracle's reference-rounding
Laplacian for it (eps 1e-2, gamma 1: the reference's own call, main.cu:222),
so that later oracle changes are caught.  data05_full.npz holds the three
full 800x552 images as raw 8-bit arrays (the input of the end-to-end CLI test).

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
sys.path.insert(0, ROOT)
import netpbm  # noqa: E402

from oracle import pyoracle as O  # noqa: E402

REF = "/root/reference/data"
X0, Y0, W, H = 16, 408, 96, 64


def image_float4(rgb_u8, maxval):
    """ppmread_float4 (ppm.h:309-330): channel / maxval in float, w = 1."""
    f = rgb_u8.astype(np.float32) / np.float32(maxval)
    return np.concatenate([f, np.ones(f.shape[:2] + (1,), np.float32)], axis=2)


def main():
    im, mv = netpbm.read(os.path.join(REF, "05.ppm"))
    sc, _ = netpbm.read(os.path.join(REF, "05_scribs.pgm"))
    gt, _ = netpbm.read(os.path.join(REF, "05_gt.pgm"))
    crop = (slice(Y0, Y0 + H), slice(X0, X0 + W))
    rgb = np.ascontiguousarray(im[crop])
    s = np.ascontiguousarray(sc[crop])
    g = np.ascontiguousarray(gt[crop])
    a, b = O.levin_scatter_u8(image_float4(rgb, mv), s, eps=1e-2, gamma=1.0)
    np.savez_compressed(os.path.join(HERE, "data05_crop.npz"), rgb=rgb, maxval=np.int32(mv), scribs=s, gt=g,
                        origin=np.array([X0, Y0], np.int32),
                        L_sha256=np.frombuffer(hashlib.sha256(a.tobytes()).digest(), np.uint8),
                        b_sha256=np.frombuffer(hashlib.sha256(b.tobytes()).digest(), np.uint8),
                        L_row_samples=a[:, ::97].copy(), b_sum=np.float64(b.astype(np.float64).sum()))
    print("wrote data05_crop.npz", rgb.shape, np.unique(s, return_counts=True))
    np.savez_compressed(os.path.join(HERE, "data05_full.npz"), rgb=im, maxval=np.int32(mv), scribs=sc, gt=gt)
    print("wrote data05_full.npz", im.shape)


if __name__ == "__main__":
    main()
