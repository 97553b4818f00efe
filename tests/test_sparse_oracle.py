"""The sparse / superpixel oracle pinned without the GPU: CSR products and the
Galerkin operator against scipy.sparse (textbook routines), the merge levels
against the reference's own loop shape, SLIC's colour conversion against
published CIELAB values and SLIC itself against a closed-form case."""
import numpy as np
import pytest
import scipy.sparse as sps

from oracle import pyoracle as O
from oracle import sporacle as SP


def _rand_coo(rng, rows, cols, n):
    i = rng.integers(0, rows, n)
    j = rng.integers(0, cols, n)
    k = rng.standard_normal(n).astype(np.float32)
    return i, j, k


def test_coo_to_csr_and_product_match_scipy():
    rng = np.random.default_rng(0)
    for rows, cols, n in ((7, 5, 40), (50, 80, 300), (3, 9, 0), (20, 20, 1)):
        i, j, k = _rand_coo(rng, rows, cols, n)
        p, jj, kk = SP.coo_to_csr(rows, cols, i, j, k)
        assert p[0] == 0 and p[-1] == n and np.all(np.diff(p) >= 0)
        # rows and columns ascending; repeats in input order
        for r in range(rows):
            seg = jj[p[r]:p[r + 1]]
            assert np.all(np.diff(seg) >= 0)
        A = sps.coo_matrix((k.astype(np.float64), (i, j)), shape=(rows, cols)).tocsr()
        x = rng.standard_normal(cols).astype(np.float32)
        y = rng.standard_normal(rows).astype(np.float32)
        want = A @ x.astype(np.float64)
        got = SP.csr_ax(p, jj, kk, x)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)
        np.testing.assert_allclose(SP.csr_ax(p, jj, kk, x, y, -1.0), want - y, rtol=1e-5, atol=1e-5)
        b0 = rng.standard_normal(rows).astype(np.float32)
        np.testing.assert_allclose(SP.csr_ax(p, jj, kk, x, y, 1.0, b0), b0 + want + y, rtol=1e-5, atol=1e-5)


def test_repeats_keep_input_order():
    p, j, k = SP.coo_to_csr(2, 3, [1, 0, 1, 1], [2, 1, 0, 2], [1.0, 2.0, 3.0, 4.0])
    assert list(p) == [0, 1, 4] and list(j) == [1, 0, 2, 2] and list(k) == [2.0, 3.0, 1.0, 4.0]


def test_restriction_and_prolongation_are_transposes():
    rng = np.random.default_rng(1)
    labels = rng.integers(0, 9, (6, 7))
    p, j, k = SP.restriction(labels, 10)  # label 9 empty
    P = sps.csr_matrix((k, j, p), shape=(10, labels.size))
    want = sps.csr_matrix((np.ones(labels.size), (labels.ravel(), np.arange(labels.size))), shape=(10, labels.size))
    assert (P != want).nnz == 0
    xc = rng.standard_normal(10).astype(np.float32)
    np.testing.assert_array_equal(SP.prolong(xc, labels, 10), (P.T @ xc).astype(np.float32))
    bad = labels.copy()
    bad[0, 0] = 10
    assert np.isnan(SP.prolong(xc, bad, 10)[0])


@pytest.mark.parametrize("W,H", [(9, 7), (12, 5)])
def test_galerkin_matches_scipy_triple_product(W, H):
    rng = np.random.default_rng(W)
    image = rng.random((H, W, 3)).astype(np.float32)
    a, _, _ = O.levin_gather_acc_f32(image, None, None, 1e-4)
    bands = O.bands25(W)
    labels = (np.arange(H)[:, None] // 3) * ((W + 3) // 4) + (np.arange(W)[None, :] // 4)
    nl = int(labels.max()) + 1
    p, j, k = SP.galerkin(a, bands, labels, nl)
    Ac = sps.csr_matrix((k.astype(np.float64), j, p), shape=(nl, nl))
    N = W * H
    rows = []
    for kk, off in enumerate(bands):
        for i in range(N):
            jn = i + off
            if 0 <= jn < N and a[kk, i] != 0:
                rows.append((i, jn, float(a[kk, i])))
    A = sps.coo_matrix(([v for _, _, v in rows], ([r for r, _, _ in rows], [c for _, c, _ in rows])), shape=(N, N))
    P = sps.csr_matrix((np.ones(N), (labels.ravel(), np.arange(N))), shape=(nl, N))
    want = (P @ A @ P.T).toarray()
    np.testing.assert_allclose(Ac.toarray(), want, rtol=1e-6, atol=1e-6)
    # the Laplacian's rows sum to 0 (no data term), so do the coarse operator's
    np.testing.assert_allclose(Ac.toarray().sum(axis=1), 0, atol=1e-4)
    # symmetric, like A
    np.testing.assert_allclose(Ac.toarray(), Ac.toarray().T, rtol=1e-5, atol=1e-5)


def test_merge_levels_sorted_equals_reference_loop():
    rng = np.random.default_rng(5)
    for trial in range(4):
        H, W = 12, 14
        labels = (np.arange(H)[:, None] // 3) * 5 + (np.arange(W)[None, :] // 3)
        nl = int(labels.max()) + 1
        # a few colours so that some means are close, some equal (ties)
        pal = rng.random((4, 3)).astype(np.float32)
        im = pal[rng.integers(0, 4, nl)][labels] + (rng.random((H, W, 3)) * 0.02 * trial).astype(np.float32)
        for nlev in (2, 3):
            a, na = SP.merge_levels(labels, nl, im, nlev, 0.0111)
            b, nb = SP.merge_levels(labels, nl, im, nlev, 0.0111, brute=True)
            assert na == nb and np.array_equal(a, b)
            assert na <= nl
        # merged pairs are adjacent superpixels of the previous level
        a, na = SP.merge_levels(labels, nl, im, 2, 0.0111)
        for c in range(na):
            olds = np.unique(labels[a == c])
            assert len(olds) in (1, 2)
            if len(olds) == 2:
                assert (min(olds), max(olds)) in SP._adjacent_pairs(labels)


def test_lab_known_values():
    # published CIELAB (D65) of sRGB white, black, red, green, blue
    im = np.array([[[1, 1, 1], [0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]], np.float32)
    lab = SP.rgb_to_lab(im)[:, 0, :].T
    want = np.array([[100, 0, 0], [0, 0, 0], [53.24, 80.09, 67.20], [87.73, -86.18, 83.18], [32.30, 79.19, -107.86]])
    np.testing.assert_allclose(lab, want, atol=0.02)


def test_slic_recovers_flat_blocks():
    """Flat colour blocks aligned with the seed grid: every block is one
    superpixel (closed form: a seed's own block is at colour distance 0)."""
    rng = np.random.default_rng(2)
    S, nbx, nby = 6, 5, 4
    W, H = S * nbx, S * nby
    pal = rng.random((nbx * nby, 3)).astype(np.float32) * 0.8 + 0.1
    blocks = (np.arange(H)[:, None] // S) * nbx + (np.arange(W)[None, :] // S)
    im = pal[blocks]
    labels, n = SP.slic(im, nbx * nby, 10.0, 10)
    assert n == nbx * nby
    assert np.array_equal(labels, blocks)  # blocks are numbered in raster order of first appearance too


def test_slic_connectivity_invariants():
    rng = np.random.default_rng(7)
    H, W = 30, 41
    im = rng.random((H, W, 3)).astype(np.float32)
    labels, n = SP.slic(im, 12, 10.0, 5)
    assert labels.min() == 0 and labels.max() == n - 1
    # numbered by first appearance
    _, first = np.unique(labels.ravel(), return_index=True)
    assert np.all(np.diff(first) > 0)
    # each label is one 4-connected piece
    relabelled, n2 = SP.connect(labels, 1)
    assert n2 == n and np.array_equal(relabelled, labels)
