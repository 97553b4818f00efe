"""Seeded synthetic matting inputs.

The reference ships one real example (data/05.ppm, 05_scribs.pgm, 05_gt.pgm)
and no generator; the benchmark workloads are synthetic images of known
ground truth built here:

    I = alpha_gt * F + (1 - alpha_gt) * B + N(0, noise^2)     (per channel, clipped to [0,1])

F and B are independent smooth colour fields (bilinear upsampling of an 8x6
grid of uniform random colours), alpha_gt is the union of soft-edged ellipses
(logistic function of the signed distance to the ellipse boundary; 10-90 %
transition ~4.4 px), and the constraints are a uniform sample of pixels whose
alpha_gt is < 0.01 or > 0.99 with weight `lam` and target alpha_gt there.

Everything is numpy with an explicit PCG64 seed, so every machine regenerates
bit-identical inputs.  Arrays are returned C-contiguous float32:
image (H, W, 3), lam (H, W), target (H, W), alpha_gt (H, W).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class MattingProblem:
    image: np.ndarray      # (H, W, 3) float32 in [0, 1]
    lam: np.ndarray        # (H, W) float32: constraint weight (0 = unconstrained)
    target: np.ndarray     # (H, W) float32: alpha target where lam > 0, else 0
    alpha_gt: np.ndarray   # (H, W) float32
    eps: float = 1e-7

    @property
    def W(self) -> int:
        return self.image.shape[1]

    @property
    def H(self) -> int:
        return self.image.shape[0]

    @property
    def N(self) -> int:
        return self.W * self.H

    def scribbles_u8(self) -> np.ndarray:
        """Scribbles in the reference's PGM convention: 255 = foreground,
        0 = background, 128 = unknown (main.cu:147, Matting.cu:205-216)."""
        s = np.full(self.lam.shape, 128, np.uint8)
        c = self.lam > 0
        s[c & (self.target >= 0.5)] = 255
        s[c & (self.target < 0.5)] = 0
        return s


def _bilinear_field(rng: np.random.Generator, W: int, H: int, gw: int = 8, gh: int = 6) -> np.ndarray:
    grid = rng.random((gh, gw, 3))
    ys = np.linspace(0.0, gh - 1.0, H)
    xs = np.linspace(0.0, gw - 1.0, W)
    y0 = np.minimum(np.floor(ys).astype(np.int64), gh - 2)
    x0 = np.minimum(np.floor(xs).astype(np.int64), gw - 2)
    fy = (ys - y0)[:, None, None]
    fx = (xs - x0)[None, :, None]
    # (1-fy)*((1-fx)*g00 + fx*g01) + fy*((1-fx)*g10 + fx*g11), with the inner
    # x interpolation done once per grid row (same operations per element)
    rows = (1 - fx) * grid[:, x0] + fx * grid[:, x0 + 1]  # (gh, W, 3)
    return (1 - fy) * rows[y0] + fy * rows[y0 + 1]


def _ellipses_alpha(rng: np.random.Generator, W: int, H: int, n: int, width: float) -> np.ndarray:
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    alpha = np.zeros((H, W))
    m = float(min(W, H))
    for _ in range(n):
        cx = rng.uniform(0.25, 0.75) * W
        cy = rng.uniform(0.25, 0.75) * H
        ra = rng.uniform(0.15, 0.35) * m
        rb = rng.uniform(0.15, 0.35) * m
        th = rng.uniform(0.0, np.pi)
        c, s = np.cos(th), np.sin(th)
        u = (xx - cx) * c + (yy - cy) * s
        v = -(xx - cx) * s + (yy - cy) * c
        # signed distance (positive inside), first-order accurate near the boundary
        rho = np.sqrt((u / ra) ** 2 + (v / rb) ** 2)
        grad = np.sqrt((u / ra**2) ** 2 + (v / rb**2) ** 2) / np.maximum(rho, 1e-12)
        d = (1.0 - rho) / np.maximum(grad, 1e-12)
        a = 1.0 / (1.0 + np.exp(-np.clip(4.0 * d / width, -60, 60)))
        alpha = np.maximum(alpha, a)
    return alpha


def make_problem(W: int, H: int, seed: int, n_ellipses: int = 1, density: float = 0.08,
                 lam: float = 100.0, eps: float = 1e-7, noise: float = 0.01,
                 width: float = 4.0) -> MattingProblem:
    """One synthetic matting problem (see module docstring)."""
    rng = np.random.default_rng(seed)
    F = _bilinear_field(rng, W, H)
    B = _bilinear_field(rng, W, H)
    a = _ellipses_alpha(rng, W, H, n_ellipses, width)
    img = a[:, :, None] * F + (1.0 - a[:, :, None]) * B + rng.normal(0.0, noise, (H, W, 3))
    img = np.clip(img, 0.0, 1.0).astype(np.float32)
    alpha_gt = a.astype(np.float32)
    cand = np.flatnonzero((alpha_gt.ravel() < 0.01) | (alpha_gt.ravel() > 0.99))
    k = min(int(round(density * W * H)), cand.size)
    pick = rng.choice(cand, size=k, replace=False) if k > 0 else np.zeros(0, np.int64)
    lam_a = np.zeros(W * H, np.float32)
    tgt = np.zeros(W * H, np.float32)
    lam_a[pick] = np.float32(lam)
    tgt[pick] = alpha_gt.ravel()[pick]
    return MattingProblem(np.ascontiguousarray(img), lam_a.reshape(H, W), tgt.reshape(H, W),
                          alpha_gt, eps)


def extra_rhs(p: MattingProblem, seed: int, density: float, lam: float = 100.0):
    """An independent scribble set on the same image (batched multi-RHS workload):
    returns (lam, target) arrays of shape (H, W)."""
    rng = np.random.default_rng(seed)
    ag = p.alpha_gt.ravel()
    cand = np.flatnonzero((ag < 0.01) | (ag > 0.99))
    k = min(int(round(density * p.N)), cand.size)
    pick = rng.choice(cand, size=k, replace=False) if k > 0 else np.zeros(0, np.int64)
    la = np.zeros(p.N, np.float32)
    tg = np.zeros(p.N, np.float32)
    la[pick] = np.float32(lam)
    tg[pick] = ag[pick]
    return la.reshape(p.H, p.W), tg.reshape(p.H, p.W)


# The workloads the benchmark and the parity tests use (sizes, seeds).
def config_small(seed: int = 1000) -> MattingProblem:
    """256x192, one ellipse, 8 % constraints, eps 1e-7, lambda 100."""
    return make_problem(256, 192, seed, 1, 0.08)


def config_1k(seed: int = 2000) -> MattingProblem:
    """1024x1024, one ellipse, 8 % constraints."""
    return make_problem(1024, 1024, seed, 1, 0.08)


def config_4k(seed: int = 3000) -> MattingProblem:
    """3840x2160, three ellipses, 4 % constraints."""
    return make_problem(3840, 2160, seed, 3, 0.04)


def config_grad(n_images: int = 16, seed: int = 5000, W: int = 512, H: int = 512):
    """Differentiation workload: n_images W x H problems, one ellipse, 8 %
    constraints each (seed + image index)."""
    return [make_problem(W, H, seed + i, 1, 0.08) for i in range(n_images)]


def config_batched_plan(n_images: int = 64, nrhs: int = 4, seed: int = 4000, wmin=240, wmax=1080, hmin=320,
                        hmax=1920):
    """The draws of config_batched without building any image: a list of
    (W, H, [density of each constraint set]) per image."""
    rng = np.random.default_rng(seed)
    plan = []
    for _ in range(n_images):
        W = int(rng.integers(wmin, wmax + 1))
        H = int(rng.integers(hmin, hmax + 1))
        plan.append((W, H, [float(rng.uniform(0.02, 0.10)) for _ in range(nrhs)]))
    return plan


def batched_image(i: int, W: int, H: int, densities, seed: int = 4000):
    """Image i of a batched workload (seeded as config_batched does): the
    problem and its constraint sets [(lam, tgt), ...]."""
    nrhs = len(densities)
    p = make_problem(W, H, seed + 1 + i, 1, 0.0)
    return p, [extra_rhs(p, seed * 100 + i * nrhs + j, d) for j, d in enumerate(densities)]


def config_batched(n_images: int = 64, nrhs: int = 4, seed: int = 4000, wmin=240, wmax=1080, hmin=320, hmax=1920,
                   pick=None):
    """Batched workload: n_images images with W ~ U[wmin,wmax], H ~ U[hmin,hmax]
    (one ellipse each), and nrhs independent constraint sets per image with
    densities ~ U[2 %, 10 %].  Returns a list of (problem, [(lam, tgt), ...]);
    with `pick` (image indices) only those images, in that order, each exactly
    as in the full workload."""
    plan = config_batched_plan(n_images, nrhs, seed, wmin, wmax, hmin, hmax)
    idx = range(n_images) if pick is None else pick
    return [batched_image(i, *plan[i], seed=seed) for i in idx]
