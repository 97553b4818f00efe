"""The library's host-side superpixel merging (gm_sp_merge_levels, no device
work) against the oracle: identical labels, level by level."""
import numpy as np
import pytest

from gpumatting_amd import sparse
from gpumatting_amd._lib import GMError
from oracle import sporacle as SP


@pytest.mark.parametrize("seed,nlev", [(0, 2), (1, 3), (2, 4)])
def test_merge_levels_matches_oracle(seed, nlev):
    rng = np.random.default_rng(seed)
    H, W = 40, 52
    labels = (np.arange(H)[:, None] // 4) * 13 + (np.arange(W)[None, :] // 4)
    nl = int(labels.max()) + 1
    pal = rng.random((6, 3)).astype(np.float32)
    im = pal[rng.integers(0, 6, nl)][labels] + (0.01 * rng.random((H, W, 3))).astype(np.float32)
    im = np.concatenate([im, np.ones((H, W, 1), np.float32)], axis=2)  # RGBA input
    got, n = sparse.merge_levels(labels, nl, im, nlev)
    want, nw = SP.merge_levels(labels, nl, im[..., :3], nlev)
    assert n == nw and np.array_equal(got, want)
    assert n < nl


def test_merge_levels_rejects_bad_labels():
    labels = np.zeros((4, 4), np.int32)
    labels[0, 0] = 3
    with pytest.raises(GMError):
        sparse.merge_levels(labels, 3, np.zeros((4, 4, 3), np.float32), 2)
