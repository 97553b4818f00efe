"""Batched multi-image, multi-RHS solver vs the FP64 oracle per system."""
import numpy as np
import pytest

from gpumatting_amd import PREC_BF16, PREC_FP16, PREC_FP32, synth
from gpumatting_amd.batch import Batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def _sparse(a64, W, N):
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for k, o in enumerate(O.bands25(W)):
        i = np.arange(N)
        m = (i + o >= 0) & (i + o < N)
        rows.append(i[m]), cols.append(i[m] + o), vals.append(a64[k][:N][m])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


@pytest.fixture(scope="module")
def small_work():
    """5 images of at most 130x90, 3 right-hand sides, with the FP64 matrix,
    FP64 solution and smallest eigenvalue of every system."""
    import scipy.sparse.linalg as sla
    work = synth.config_batched(n_images=5, nrhs=3, seed=77, wmin=20, wmax=130, hmin=15, hmax=90)
    ref = {}
    for i, (p, rhs) in enumerate(work):
        for j, (lam, tgt) in enumerate(rhs):
            a, bb, _ = O.levin_gather_acc_f32(p.image, lam, tgt, 1e-7)
            a64, b64 = a.astype(np.float64), bb.astype(np.float64)
            xe, _, _ = O.pcg_f64(a64, O.bands25(p.W), b64, np.zeros(p.N), O.pad_for(p.W), 1e-13)
            A = _sparse(a64, p.W, p.N)
            lmin = sla.eigsh(A, k=1, sigma=0, which="LM", return_eigenvectors=False)[0]
            assert lmin > 0
            ref[i, j] = (A, b64, xe, lmin)
    return work, ref


# per_cu 16 (the default: at most one tile per workgroup on this workload);
# 1 and 2 put several tiles and several images in one workgroup's range, so the
# LDS ring carry and the multi-image slot index run; tile widths 64 and 256
# beside the default 128.
@pytest.mark.parametrize("prec,per_cu,tile_w", [(PREC_FP32, None, None), (PREC_FP16, None, None),
                                                (PREC_BF16, None, None), (PREC_FP32, 1, None), (PREC_FP16, 1, 60),
                                                (PREC_FP32, 2, 252), (PREC_BF16, 2, None), (PREC_FP16, None, 252),
                                                (PREC_FP32, None, 60)])
def test_batched_matches_fp64_per_system(gpu, small_work, prec, per_cu, tile_w):
    """Every system reaches the tolerance, checked twice: by the solver's own
    FP64 true residual and by an independent FP64 residual of the returned
    alpha on the host; the alpha error against the FP64 solution is within
    what that residual allows (||x - x*|| <= ||A x - b|| / lambda_min(A))."""
    work, ref = small_work
    b = Batch([(p.W, p.H) for p, _ in work], 3, prec)
    if per_cu is not None:
        b.set_grid(per_cu)
    if tile_w is not None:
        b.set_tiles(tile_w)
    for i, (p, rhs) in enumerate(work):
        b.assemble(i, p.image, 1e-7)
        for j, (lam, tgt) in enumerate(rhs):
            b.set_rhs(i, j, lam, tgt)
    b.solve(tol=1e-7, max_iter=20000, check_every=5)
    for i, (p, rhs) in enumerate(work):
        for j, (lam, tgt) in enumerate(rhs):
            inf = b.info(i, j)
            assert inf.status == 0 and inf.true_relres <= 1e-7, (i, j, inf.true_relres)
            A, b64, xe, lmin = ref[i, j]
            x = b.alpha(i, j).cpu().numpy().astype(np.float64)
            res = np.linalg.norm(A @ x - b64)
            assert res <= 1.01e-7 * np.linalg.norm(b64), (i, j, res)
            err = np.linalg.norm(x - xe)
            assert err <= res / lmin + 1e-6 * np.linalg.norm(xe), (i, j, err, res / lmin)
            assert np.abs(x - xe).max() <= 1e-2, (i, j)
    b.close()


def test_batched_partition_covers_every_tile_once(gpu, small_work):
    """The same batch solved at several grids and tile widths gives the same
    iterates up to the order of the FP64 partial sums: equal iteration counts
    (within 2) and alphas within the tolerance's reach."""
    work, _ = small_work
    outs = []
    for per_cu, tw in ((16, 124), (1, 124), (2, 124), (16, 60), (1, 252), (4, 252)):
        b = Batch([(p.W, p.H) for p, _ in work], 3, PREC_FP32)
        b.set_grid(per_cu)
        b.set_tiles(tw)
        for i, (p, rhs) in enumerate(work):
            b.assemble(i, p.image, 1e-7)
            for j, (lam, tgt) in enumerate(rhs):
                b.set_rhs(i, j, lam, tgt)
        b.solve(tol=1e-7, max_iter=20000, check_every=5)
        outs.append({(i, j): (b.info(i, j).iterations, b.alpha(i, j).cpu().numpy())
                     for i in range(len(work)) for j in range(3)})
        b.close()
    for o in outs[1:]:
        for key, (it, x) in o.items():
            it0, x0 = outs[0][key]
            assert abs(it - it0) <= 2, (key, it, it0)
            assert np.abs(x - x0).max() <= 1e-3, key


def test_batched_tolerance_and_iterations_match_single_solver(gpu):
    from gpumatting_amd import matting, METHOD_PCG
    work = synth.config_batched(n_images=3, nrhs=2, seed=5, wmin=60, wmax=200, hmin=50, hmax=160)
    b = Batch([(p.W, p.H) for p, _ in work], 2, PREC_FP32)
    for i, (p, rhs) in enumerate(work):
        b.assemble(i, p.image, 1e-7)
        for j, (lam, tgt) in enumerate(rhs):
            b.set_rhs(i, j, lam, tgt)
    b.solve(tol=1e-6, check_every=4)
    for i, (p, rhs) in enumerate(work):
        for j, (lam, tgt) in enumerate(rhs):
            s = matting.assemble(p.image, lam, tgt, 1e-7)
            x1, info1 = matting.solve(s, method=METHOD_PCG, tol=1e-6)
            inf = b.info(i, j)
            assert inf.true_relres <= 1e-6
            assert abs(inf.iterations - info1.iterations) <= max(6, 0.05 * info1.iterations)
    b.close()


def test_batched_autotune_then_solve(gpu):
    """The tuner times every tile width and grid on the batch itself, keeps a
    candidate, leaves the initial guesses as they were, and the solve that
    follows converges every system."""
    import torch
    work = synth.config_batched(n_images=4, nrhs=2, seed=11, wmin=40, wmax=300, hmin=30, hmax=200)
    b = Batch([(p.W, p.H) for p, _ in work], 2, PREC_FP16)
    rng = np.random.default_rng(3)
    x0s = {}
    for i, (p, rhs) in enumerate(work):
        b.assemble(i, p.image, 1e-7)
        for j, (lam, tgt) in enumerate(rhs):
            x0 = rng.random(p.N).astype(np.float32)
            x0s[i, j] = x0
            b.set_rhs(i, j, lam, tgt, x0)
    for solve in (None, dict(tol=1e-6, check_every=6)):  # launch timing, then whole-solve timing
        us = b.autotune(reps=2, solve=solve)
        tw, pc, tuned = b.config()
        assert tw in (60, 124, 252) and pc in (4, 8, 16, 32) and us > 0 and tuned == us
        torch.cuda.synchronize()
        for (i, j), x0 in x0s.items():
            assert np.array_equal(b.alpha(i, j).cpu().numpy(), x0)
    b.solve(tol=1e-6, check_every=6)
    for i, (p, rhs) in enumerate(work):
        for j in range(2):
            inf = b.info(i, j)
            assert inf.status == 0 and inf.true_relres <= 1e-6, (i, j, inf.true_relres)
    b.close()
