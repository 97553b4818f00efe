"""CPU oracle of the sparse-matrix and superpixel operators -- TEST INFRASTRUCTURE.

Only tests/ import this.  numpy restatements, for small inputs, of what
include/gpumatting.h specifies for:
  * CSR / COO products and the COO -> CSR conversion (CompressedMatrix.cu:45-165,
    CoordMatrix.cu:50-87, with the evident fixes: element counts, row bound,
    y[row] instead of y[0], a fixed summation order instead of atomics);
  * the superpixel restriction P (gridoperators.h:41-57), the prolongation P^T
    and the Galerkin coarse operator P A P^T of a banded matrix;
  * the merging levels of downsampleOps (gridoperators.h:59-138, its open
    points filled in as documented);
  * SLIC (Achanta et al., TPAMI 2012, Algorithm 1, as specified for gm_slic).
Float32 values where the library rounds to float32, float64 sums where it
accumulates in FP64.
"""
from __future__ import annotations

import numpy as np

F = np.float32


# ---------------------------------------------------------------------------
# CSR / COO

def csr_ax(p, j, k, x, y=None, ysign=1.0, b0=None):
    """Row r: float32(sum in entry order of float64(float32(k_e * x_j))),
    then (+ b0[r]) and (+ ysign * y[r]) in float32."""
    p, j = np.asarray(p, np.int64), np.asarray(j, np.int64)
    prod = (np.asarray(k, F) * np.asarray(x, F)[j]).astype(np.float64)
    rows = len(p) - 1
    out = np.empty(rows, F)
    for r in range(rows):
        s = 0.0
        for e in range(p[r], p[r + 1]):
            s += prod[e]
        v = F(s)
        if b0 is not None:
            v = F(b0[r] + v)
        if y is not None:
            v = F(v + F(ysign) * F(y[r]))
        out[r] = v
    return out


def coo_to_csr(rows, cols, i, j, k):
    """Entries stably sorted by (row, column); repeats keep their order."""
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    key = i * max(cols, 1) + j
    order = np.argsort(key, kind="stable")
    p = np.searchsorted(key[order], np.arange(rows + 1, dtype=np.int64) * max(cols, 1), side="left")
    return p.astype(np.int32), j[order].astype(np.int32), np.asarray(k, F)[order]


# ---------------------------------------------------------------------------
# superpixel operators

def restriction(labels, nlabels):
    """P[labels[u], u] = 1, each row's pixels ascending."""
    lab = np.asarray(labels, np.int64).ravel()
    order = np.argsort(lab, kind="stable")
    p = np.searchsorted(lab[order], np.arange(nlabels + 1), side="left")
    return p.astype(np.int32), order.astype(np.int32), np.ones(len(lab), F)


def galerkin(a, bands, labels, nlabels):
    """P A P^T: one entry per (row, column), the fine entries (pixel i, band k)
    with a neighbour inside the matrix and a nonzero value summed in float64
    in (i, k) order, rounded to float32."""
    a = np.asarray(a, F)
    nb = len(bands)
    lab = np.asarray(labels, np.int64).ravel()
    N = len(lab)
    i = np.repeat(np.arange(N, dtype=np.int64), nb)
    kb = np.tile(np.arange(nb, dtype=np.int64), N)
    jn = i + np.asarray(bands, np.int64)[kb]
    val = a[kb, i] if len(i) else np.zeros(0, F)
    ok = (jn >= 0) & (jn < N) & (val != 0)
    key = lab[i[ok]] * nlabels + lab[jn[ok]]
    v = val[ok].astype(np.float64)
    order = np.argsort(key, kind="stable")
    ks, vs = key[order], v[order]
    if len(ks) == 0:
        return np.zeros(nlabels + 1, np.int32), np.zeros(0, np.int32), np.zeros(0, F)
    head = np.r_[True, ks[1:] != ks[:-1]]
    starts = np.flatnonzero(head)
    sums = np.empty(len(starts), np.float64)
    ends = np.r_[starts[1:], len(ks)]
    for q, (s0, s1) in enumerate(zip(starts, ends)):
        t = 0.0
        for e in range(s0, s1):
            t += vs[e]
        sums[q] = t
    rk = ks[starts]
    p = np.searchsorted(rk, np.arange(nlabels + 1, dtype=np.int64) * nlabels, side="left")
    return p.astype(np.int32), (rk % nlabels).astype(np.int32), sums.astype(F)


def prolong(xc, labels, nlabels):
    lab = np.asarray(labels, np.int64).ravel()
    ok = (lab >= 0) & (lab < nlabels)
    out = np.full(len(lab), np.nan, F)
    out[ok] = np.asarray(xc, F)[lab[ok]]
    return out


def _means_f32(labels, nl, im):
    H, W = labels.shape
    means = np.zeros((nl, 3), F)
    size = np.zeros(nl, np.int64)
    for y in range(H):
        for x in range(W):
            c = labels[y, x]
            for q in range(3):
                means[c, q] = F(means[c, q] + F(im[y, x, q]))
            size[c] += 1
    for c in range(nl):
        if size[c]:
            means[c] = (means[c] / F(size[c])).astype(F)
    return means


def _adjacent_pairs(labels):
    a = labels
    pairs = set()
    h = a[:, :-1] != a[:, 1:]
    for c, d in zip(a[:, :-1][h], a[:, 1:][h]):
        pairs.add((min(c, d), max(c, d)))
    v = a[:-1, :] != a[1:, :]
    for c, d in zip(a[:-1, :][v], a[1:, :][v]):
        pairs.add((min(c, d), max(c, d)))
    return sorted(pairs)


def _dist_f32(mu, mv):
    t = F(mu[0] - mv[0])
    t = F(t * t)
    d = t
    t = F(mu[1] - mv[1])
    t = F(t * t)
    d = F(d + t)
    t = F(mu[2] - mv[2])
    t = F(t * t)
    return F(d + t)


def merge_levels(labels, nlabels, im, nlevels, threshold=0.0111, brute=False):
    """Levels 2..nlevels.  brute=True runs the reference's own loop shape
    (gridoperators.h:96-124: repeated scan for the first minimum over unmerged
    adjacent pairs, u then v ascending); otherwise one ascending sort."""
    lab = np.array(labels, np.int64).copy()
    nl = nlabels
    im = np.asarray(im, F)
    thr = F(threshold)
    for _ in range(1, nlevels):
        means = _means_f32(lab, nl, im)
        pairs = _adjacent_pairs(lab)
        dist = {pr: _dist_f32(means[pr[0]], means[pr[1]]) for pr in pairs}
        merged = np.zeros(nl, bool)
        mapping = np.full(nl, -1, np.int64)
        cur = 0
        if brute:
            while True:
                best, bu, bv = F(1e6), -1, -1
                for (u, v) in pairs:
                    if merged[u] or merged[v]:
                        continue
                    if dist[(u, v)] < best:
                        best, bu, bv = dist[(u, v)], u, v
                if best > thr or bu < 0:
                    break
                merged[bu] = merged[bv] = True
                mapping[bu] = mapping[bv] = cur
                cur += 1
        else:
            for (u, v) in sorted(pairs, key=lambda pr: (dist[pr], pr[0], pr[1])):
                if not (dist[(u, v)] <= thr):
                    break
                if merged[u] or merged[v]:
                    continue
                merged[u] = merged[v] = True
                mapping[u] = mapping[v] = cur
                cur += 1
        for c in range(nl):
            if not merged[c]:
                mapping[c] = cur
                cur += 1
        lab = mapping[lab]
        nl = cur
    return lab.astype(np.int32), nl


# ---------------------------------------------------------------------------
# SLIC

def rgb_to_lab(im):
    """(H, W, >=3) RGB in [0, 1] -> (3, H, W) float32 CIELAB (sRGB, D65),
    evaluated in FP64 and rounded to the nearest multiple of 2^-16 (half to
    even), which float32 holds exactly for |value| < 128."""
    c = np.asarray(im, np.float32)[..., :3].astype(np.float64)
    c = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    X = (0.4124564 * c[..., 0] + 0.3575761 * c[..., 1] + 0.1804375 * c[..., 2]) / 0.95047
    Y = 0.2126729 * c[..., 0] + 0.7151522 * c[..., 1] + 0.0721750 * c[..., 2]
    Z = (0.0193339 * c[..., 0] + 0.1191920 * c[..., 1] + 0.9503041 * c[..., 2]) / 1.08883

    def f(t):
        return np.where(t > 0.008856, np.cbrt(t), 7.787 * t + 16.0 / 116.0)

    fx, fy, fz = f(X), f(Y), f(Z)
    lab = np.stack([116.0 * fy - 16.0, 500.0 * (fx - fy), 200.0 * (fy - fz)])
    return (np.rint(lab * 65536.0) / 65536.0).astype(F)


def slic_grid(W, H, K):
    N = W * H
    S = max(1, int(np.sqrt(N / K) + 0.5))
    nsx = max(1, int(W / S + 0.5))
    nsy = max(1, int(H / S + 0.5))
    return S, nsx, nsy


def _cell(v, n, extent):
    return np.minimum(n - 1, (np.asarray(v, np.int64) * n) // extent)


def slic_raw(im, K, m=10.0, iters=10, lab=None, return_seeds=False):
    """SLIC before the connectivity step: (labels (H, W), S), and the final
    seeds (5, nseeds) with return_seeds.  lab: start from these (3, H, W)
    planes instead of converting im.  Every float operation is one float32
    operation in the order written (numpy float32 scalars and arrays), so the
    result is exactly that of any evaluation with the same rounding steps."""
    if lab is None:
        im = np.asarray(im, F)
        lab = rgb_to_lab(im)
    lab = np.asarray(lab, F)
    H, W = lab.shape[1:]
    S, nsx, nsy = slic_grid(W, H, K)
    ns = nsx * nsy
    seeds = np.zeros((5, ns), F)
    for k in range(ns):
        iy, ix = divmod(k, nsx)
        sx = int(F(F(ix + 0.5) * F(W)) / F(nsx))
        sy = int(F(F(iy + 0.5) * F(H)) / F(nsy))
        sx, sy = min(max(sx, 0), W - 1), min(max(sy, 0), H - 1)
        bx, by, best = sx, sy, np.inf
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                x, y = sx + dx, sy + dy
                if x < 1 or x > W - 2 or y < 1 or y > H - 2:
                    continue
                g = F(0)
                for q in range(3):
                    ddx = F(lab[q, y, x + 1] - lab[q, y, x - 1])
                    ddy = F(lab[q, y + 1, x] - lab[q, y - 1, x])
                    g = F(g + F(F(ddx * ddx) + F(ddy * ddy)))
                if g < best:
                    best, bx, by = g, x, y
        seeds[:, k] = (lab[0, by, bx], lab[1, by, bx], lab[2, by, bx], bx, by)
    yy, xx = np.mgrid[0:H, 0:W]
    cx, cy = _cell(xx, nsx, W), _cell(yy, nsy, H)
    labels = cy * nsx + cx
    wd = F(F(m) * F(m) / F(S * S))
    Sf = F(S)
    fx, fy = xx.astype(F), yy.astype(F)
    for _ in range(iters):
        best = np.full((H, W), np.inf, F)
        bk = np.full((H, W), -1, np.int64)
        for oy in range(-2, 3):
            for ox in range(-2, 3):
                jy, jx = cy + oy, cx + ox
                inside = (jy >= 0) & (jy < nsy) & (jx >= 0) & (jx < nsx)
                k = np.where(inside, jy * nsx + jx, 0)
                ex = (fx - seeds[3, k]).astype(F)
                ey = (fy - seeds[4, k]).astype(F)
                ok = inside & (np.abs(ex) <= Sf) & (np.abs(ey) <= Sf)
                dl = lab[0] - seeds[0, k]
                da = lab[1] - seeds[1, k]
                db = lab[2] - seeds[2, k]
                D = ((dl * dl + da * da + db * db) + (ex * ex + ey * ey) * wd).astype(F)
                take = ok & (D < best)
                best = np.where(take, D, best)
                bk = np.where(take, k, bk)
        labels = np.where(bk >= 0, bk, labels)
        new = seeds.copy()
        lk = labels.ravel()
        X, Y = seeds[3, lk], seeds[4, lk]
        inwin = (np.abs(fx.ravel() - X) <= Sf) & (np.abs(fy.ravel() - Y) <= Sf)
        cnt = np.bincount(lk[inwin], minlength=ns).astype(np.float64)
        feats = [lab[0].ravel(), lab[1].ravel(), lab[2].ravel(), fx.ravel(), fy.ravel()]
        for q in range(5):
            t = np.bincount(lk[inwin], weights=feats[q][inwin].astype(np.float64), minlength=ns)
            new[q] = np.where(cnt > 0, (t / np.maximum(cnt, 1)).astype(F), seeds[q])
        seeds = new
    if return_seeds:
        return labels.astype(np.int32), S, seeds
    return labels.astype(np.int32), S


def connect(labels, minsize):
    """Connected pieces, small ones joined to the left (else upper) neighbour
    of their first pixel; numbering in raster order of first appearance."""
    H, W = labels.shape
    lab = labels.ravel()
    N = H * W
    out = np.full(N, -1, np.int64)
    nxt = 0
    for u0 in range(N):
        if out[u0] >= 0:
            continue
        c = lab[u0]
        piece, stack = [], [u0]
        out[u0] = -2
        while stack:
            u = stack.pop()
            piece.append(u)
            y, x = divmod(u, W)
            for v, okv in ((u - 1, x > 0), (u + 1, x + 1 < W), (u - W, y > 0), (u + W, y + 1 < H)):
                if okv and out[v] == -1 and lab[v] == c:
                    out[v] = -2
                    stack.append(v)
        y0, x0 = divmod(u0, W)
        if len(piece) < minsize and (x0 > 0 or y0 > 0):
            tgt = out[u0 - 1] if x0 > 0 else out[u0 - W]
        else:
            tgt = nxt
            nxt += 1
        out[piece] = tgt
    return out.reshape(H, W).astype(np.int32), nxt


def slic(im, K, m=10.0, iters=10):
    raw, S = slic_raw(im, K, m, iters)
    return connect(raw, max(1, S * S // 4))
