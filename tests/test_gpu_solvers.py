"""GPU solver parity through the C ABI: cgSolve / gradSolve at equal iteration
count vs the oracle's restatement (main.cu:406-526), Jacobi-PCG and the
16-bit mixed forms vs the FP64 oracle, smoothers vs the reference's
Solve.cpp semantics, failure statuses, determinism, and a 4K solve."""
import ctypes as C

import numpy as np
import pytest
import torch

from gpumatting_amd import (ASSEMBLY_ACCURATE, ASSEMBLY_REFERENCE, METHOD_CG, METHOD_GD, METHOD_PCG, PREC_BF16,
                            PREC_FP16, PREC_FP32, _lib, matting, synth)
from gpumatting_amd._lib import BandedMatrix, call, lib
from gpumatting_amd.matting import DeviceMatrix, Padded
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
S0 = C.c_void_p(0)


def dptr(t):
    return C.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def small():
    p = synth.config_small()
    a, b, d = O.levin_gather_acc_f32(p.image, p.lam, p.target, 1e-7)
    bands = O.bands25(p.W)
    pad = O.pad_for(p.W)
    x_exact, it, rr = O.pcg_f64(a.astype(np.float64), bands, b.astype(np.float64), np.zeros(p.N), pad, 1e-13)
    return p, a, b, bands, pad, x_exact


def _gpu_system(p, mode=ASSEMBLY_ACCURATE):
    return matting.assemble(p.image, p.lam, p.target, 1e-7, mode)


@pytest.mark.parametrize("iters,restart", [(50, 101), (200, 101), (150, 40)])
def test_cgsolve_matches_reference_recurrence(gpu, small, iters, restart):
    p, a, b, bands, pad, _ = small
    s = _gpu_system(p)
    x = Padded(p.N, pad, init=np.full(p.N, 0.5, np.float32))
    call("gm_cg_solve", x.ptr, s.L.ref, dptr(s.b), pad, iters, restart, S0)
    torch.cuda.synchronize()
    xr = O.cg_ref_f32(a, bands, b, np.full(p.N, 0.5, np.float32), pad, iters, restart)
    assert np.abs(x.numpy() - xr).max() <= 1e-3


def test_gradsolve_matches_reference_recurrence(gpu, small):
    p, a, b, bands, pad, _ = small
    s = _gpu_system(p)
    x = Padded(p.N, pad, init=np.full(p.N, 0.5, np.float32))
    call("gm_grad_solve", x.ptr, s.L.ref, dptr(s.b), 60, pad, S0)
    torch.cuda.synchronize()
    xr = O.gd_ref_f32(a, bands, b, np.full(p.N, 0.5, np.float32), pad, 60)
    assert np.abs(x.numpy() - xr).max() <= 1e-3


@pytest.fixture(scope="module")
def k1024():
    p = synth.config_1k()
    a, b, d = O.levin_gather_acc_f32(p.image, p.lam, p.target, 1e-7)
    return p, a, b, O.bands25(p.W), O.pad_for(p.W)


@pytest.mark.parametrize("iters,restart", [(200, 101), (120, 40)])
def test_cgsolve_matches_reference_recurrence_1024(gpu, k1024, iters, restart):
    """cgSolve at equal iteration count on the 1024x1024 workload."""
    p, a, b, bands, pad = k1024
    s = _gpu_system(p)
    x = Padded(p.N, pad, init=np.full(p.N, 0.5, np.float32))
    call("gm_cg_solve", x.ptr, s.L.ref, dptr(s.b), pad, iters, restart, S0)
    torch.cuda.synchronize()
    xr = O.cg_ref_f32(a, bands, b, np.full(p.N, 0.5, np.float32), pad, iters, restart)
    d = np.abs(x.numpy() - xr)
    print(f"\n[1024^2] cgSolve {iters} iterations restart {restart}: max |da| {d.max():.2e}")
    assert d.max() <= 1e-3


def test_gradsolve_matches_reference_recurrence_1024(gpu, k1024):
    p, a, b, bands, pad = k1024
    s = _gpu_system(p)
    x = Padded(p.N, pad, init=np.full(p.N, 0.5, np.float32))
    call("gm_grad_solve", x.ptr, s.L.ref, dptr(s.b), 60, pad, S0)
    torch.cuda.synchronize()
    xr = O.gd_ref_f32(a, bands, b, np.full(p.N, 0.5, np.float32), pad, 60)
    d = np.abs(x.numpy() - xr)
    print(f"\n[1024^2] gradSolve 60 iterations: max |da| {d.max():.2e}")
    assert d.max() <= 1e-3


@pytest.mark.parametrize("method", [METHOD_CG, METHOD_PCG])
def test_tolerance_solve_vs_fp64(gpu, small, method):
    p, a, b, bands, pad, x_exact = small
    s = _gpu_system(p)
    x, info = matting.solve(s, method=method, tol=1e-6, max_iter=5000)
    assert info.status == 0 and info.true_relres <= 1e-6
    xo, it, _ = O.pcg_f32(a, bands, b, np.zeros(p.N, np.float32), pad, 1e-6, 5000, precond=method == METHOD_PCG)
    assert abs(info.iterations - it) <= max(5, 0.05 * it)
    assert np.abs(x.numpy() - x_exact).max() <= 5e-3
    x2, info2 = matting.solve(s, method=method, tol=5e-8, max_iter=5000)
    assert np.abs(x2.numpy() - x_exact).max() <= 2e-4
    x3, info3 = matting.solve(s, method=method, tol=1e-10, max_iter=20000, refine_mode=1)
    assert info3.status == 0 and info3.true_relres <= 1e-10
    assert np.abs(x3.numpy() - x_exact).max() <= 1e-6


@pytest.mark.parametrize("prec", [PREC_FP16, PREC_BF16])
def test_mixed_precision_vs_fp64(gpu, small, prec):
    p, a, b, bands, pad, x_exact = small
    s = _gpu_system(p)
    _, info32 = matting.solve(s, method=METHOD_PCG, tol=1e-6, max_iter=5000)
    x, info = matting.solve(s, method=METHOD_PCG, precision=prec, tol=1e-6, max_iter=5000)
    assert info.status == 0 and info.true_relres <= 1e-6
    assert info.replacements >= 1
    # FP16 keeps FP32's iteration count; BF16's 8-bit significand costs iterations
    assert info.iterations <= (1.2 if prec == PREC_FP16 else 2.5) * info32.iterations + 10
    assert np.abs(x.numpy() - x_exact).max() <= 5e-3
    if prec == PREC_FP16:
        # FP64 iterative refinement around the 16-bit inner solves reaches the FP64 answer
        x2, info2 = matting.solve(s, method=METHOD_PCG, precision=prec, tol=1e-9, max_iter=20000, refine_mode=1)
        assert info2.status == 0 and info2.true_relres <= 1e-9
        assert np.abs(x2.numpy() - x_exact).max() <= 1e-5


def test_solve_is_bitwise_deterministic(gpu, small):
    p = small[0]
    s = _gpu_system(p)
    outs = [matting.solve(s, method=METHOD_PCG, tol=1e-6)[0].numpy() for _ in range(2)]
    outs += [matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=1e-6)[0].numpy() for _ in range(2)]
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[2], outs[3])


def test_true_relres_matches_oracle(gpu, small):
    p, a, b, bands, pad, _ = small
    s = _gpu_system(p)
    x = np.random.default_rng(0).random(p.N).astype(np.float32)
    got = matting.true_relres(s.L, Padded(p.N, pad, init=x), s.b)
    assert got == pytest.approx(O.true_relres_f32(a, bands, b, x, pad), rel=1e-9)


def _tridiag(N, diag, off):
    bands = np.array([-1, 0, 1], np.int32)
    a = np.zeros((3, N), np.float32)
    a[0, 1:] = off
    a[1] = diag
    a[2, :-1] = off
    return bands, a


def test_breakdown_and_nonfinite_statuses(gpu):
    N = 1000
    bands, a = _tridiag(N, -2.0, 0.5)  # negative definite
    M = DeviceMatrix(N, bands, a)
    b = torch.ones(N, device="cuda")
    x, info = matting.solve(M, b, pad=1, method=METHOD_CG, raise_on_error=False)
    assert info.status == 4  # GM_ERR_BREAKDOWN
    bands, a = _tridiag(N, 4.0, -1.0)
    M = DeviceMatrix(N, bands, a)
    b[17] = float("nan")
    x, info = matting.solve(M, b, pad=1, method=METHOD_PCG, raise_on_error=False)
    assert info.status == 5  # GM_ERR_NONFINITE
    b[17] = 1.0
    x, info = matting.solve(M, b, pad=1, method=METHOD_PCG, tol=1e-30, max_iter=3, raise_on_error=False)
    assert info.status == 6 and info.iterations == 3  # GM_ERR_NOT_CONVERGED
    x, info = matting.solve(M, b, pad=1, method=METHOD_PCG, tol=1e-6)
    xd = np.linalg.solve(np.diag(np.full(N, 4.0)) + np.diag(np.full(N - 1, -1.0), 1) + np.diag(np.full(N - 1, -1.0), -1),
                         np.ones(N))
    assert np.abs(x.numpy() - xd).max() <= 1e-5


def test_generic_band_structure_cg(gpu):
    """Non-Levin band sets take the generic kernels; same answers."""
    N = 5000
    bands = np.array([-37, -1, 0, 1, 37], np.int32)
    a = np.zeros((5, N), np.float32)
    i = np.arange(N)
    a[0, i >= 37] = -0.5
    a[1, i >= 1] = -1
    a[2] = 4.5
    a[3, i + 1 < N] = -1
    a[4, i + 37 < N] = -0.5
    M = DeviceMatrix(N, bands, a)
    rhs = np.random.default_rng(3).random(N).astype(np.float32)
    b = torch.from_numpy(rhs).cuda()
    for method in (METHOD_CG, METHOD_PCG):
        x, info = matting.solve(M, b, pad=37, method=method, tol=1e-6)
        xo, it, _ = O.pcg_f32(a, bands, rhs, np.zeros(N, np.float32), 37, 1e-6, 5000, precond=method == METHOD_PCG)
        assert np.abs(x.numpy() - xo).max() <= 1e-4
    with pytest.raises(_lib.GMError):
        matting.solve(M, b, pad=37, method=METHOD_PCG, precision=PREC_FP16)


def test_tiny_images(gpu):
    for (W, H) in [(1, 1), (2, 5), (3, 3)]:
        p = synth.make_problem(W, H, 1, 1, 0.0)
        lam = np.full((H, W), 100.0, np.float32)
        tgt = np.random.default_rng(W).random((H, W)).astype(np.float32)
        s = matting.assemble(p.image, lam, tgt, 1e-7)
        x, info = matting.solve(s, method=METHOD_PCG, tol=1e-7)
        a, b, _ = O.levin_gather_acc_f32(p.image, lam, tgt, 1e-7)
        xe, _, _ = O.pcg_f64(a.astype(np.float64), O.bands25(W), b.astype(np.float64), np.zeros(W * H),
                             O.pad_for(W), 1e-13)
        assert np.abs(x.numpy() - xe).max() <= 1e-5


def _dd_levin(W=37, H=21):
    p = synth.make_problem(W, H, 8, 1, 0.5)
    a, b, _ = O.levin_gather_acc_f32(p.image, p.lam, p.target, 1e-2)
    a[12] += 2.0
    return p, a, b


@pytest.mark.parametrize("iters", [1, 2, 7])
def test_jacobi_host_bitwise(gpu, iters):
    p, a, b = _dd_levin()
    bands, pad = O.bands25(p.W), O.pad_for(p.W)
    x0 = np.random.default_rng(2).random(p.N).astype(np.float32)
    xp = O.padded(x0, pad)
    h = BandedMatrix(a.ctypes.data, a.shape[1], bands.ctypes.data, 25, p.N, p.N)
    call("gm_jacobi_host", xp.ctypes.data + 4 * pad, C.byref(h), b.ctypes.data, iters, pad, C.c_float(2.0 / 3.0))
    assert np.array_equal(xp[pad:pad + p.N], O.jacobi_f32(a, bands, b, x0, pad, iters))


@pytest.mark.parametrize("levin", [True, False])
def test_gauss_seidel_host_bitwise(gpu, levin):
    if levin:
        p, a, b = _dd_levin()
        bands, pad, N = O.bands25(p.W), O.pad_for(p.W), p.N
    else:
        N = 3001
        bands = np.array([-10, 0, 10], np.int32)
        a = np.zeros((3, N), np.float32)
        i = np.arange(N)
        rng = np.random.default_rng(4)
        a[0, i >= 10] = rng.random((i >= 10).sum()) - 0.5
        a[1] = 3.0
        a[2, i + 10 < N] = rng.random((i + 10 < N).sum()) - 0.5
        b = rng.random(N).astype(np.float32)
        pad = 10
    x0 = np.random.default_rng(5).random(N).astype(np.float32)
    xp = O.padded(x0, pad)
    h = BandedMatrix(a.ctypes.data, a.shape[1], bands.ctypes.data, len(bands), N, N)
    call("gm_gauss_seidel_host", xp.ctypes.data + 4 * pad, C.byref(h), b.ctypes.data, 3)
    assert np.array_equal(xp[pad:pad + N], O.gs_f32(a, bands, b, x0, pad, 3))


def test_gpu_jacobi_driver(gpu):
    p, a, b = _dd_levin()
    bands, pad = O.bands25(p.W), O.pad_for(p.W)
    M = DeviceMatrix(p.N, bands, a)
    x0 = np.random.default_rng(6).random(p.N).astype(np.float32)
    for iters in (1, 4, 9):
        x = Padded(p.N, pad, init=x0)
        bt = torch.from_numpy(b).cuda()
        call("gm_jacobi", x.ptr, M.ref, dptr(bt), C.c_float(2.0 / 3.0), pad, iters, S0)
        torch.cuda.synchronize()
        ref = O.jacobi_f32(a, bands, b, x0, pad, iters)
        assert np.abs(x.numpy() - ref).max() <= 1e-5


@pytest.mark.parametrize("method,prec", [(METHOD_CG, PREC_FP32), (METHOD_PCG, PREC_FP32), (METHOD_PCG, PREC_FP16),
                                         (METHOD_PCG, PREC_BF16)])
def test_refine_mode2_fp64_master(gpu, small, method, prec):
    """refine_mode 2 keeps x's FP64 master inside the iteration: it reaches
    tolerances below the FP32-x floor (~2e-8 here) and the FP64 answer."""
    p, a, b, bands, pad, x_exact = small
    s = _gpu_system(p)
    x, info = matting.solve(s, method=method, precision=prec, tol=1e-10, max_iter=20000, refine_mode=2)
    assert info.status == 0 and info.true_relres <= 1e-10
    assert info.replacements >= 1
    assert np.abs(x.numpy() - x_exact).max() <= 1e-6
    # an initial guess is honoured (moved into the master before the first
    # residual): the FP32-rounded answer (relres ~1e-8) needs few iterations
    x0 = Padded(p.N, pad, init=x.numpy())
    _, info2 = matting.solve(s, x=x0, method=method, precision=prec, tol=1e-10, max_iter=20000, refine_mode=2)
    assert info2.status == 0 and info2.iterations <= info.iterations // 3
    # with the periodic restart of cgSolve (r recomputed from the master, p = -z)
    x3, info3 = matting.solve(s, method=method, precision=prec, tol=1e-9, max_iter=20000, refine_mode=2,
                              restart=40)
    assert info3.status == 0 and np.abs(x3.numpy() - x_exact).max() <= 1e-5


def test_generic_bands_refine_mode2(gpu):
    """refine_mode 2 through the generic-band kernels (a band set that is not
    the 5x5 Levin structure), against the oracle's FP64 PCG on the same
    matrix."""
    N = 5000
    bands = np.array([-37, -1, 0, 1, 37], np.int32)
    a = np.zeros((5, N), np.float32)
    i = np.arange(N)
    a[0, i >= 37] = -0.5
    a[1, i >= 1] = -1
    a[2] = 3.02
    a[3, i + 1 < N] = -1
    a[4, i + 37 < N] = -0.5
    M = DeviceMatrix(N, bands, a)
    rhs = np.random.default_rng(5).random(N).astype(np.float32)
    b = torch.from_numpy(rhs).cuda()
    xe, _, _ = O.pcg_f64(a.astype(np.float64), bands, rhs.astype(np.float64), np.zeros(N), 37, 1e-14)
    for method in (METHOD_CG, METHOD_PCG):
        x, info = matting.solve(M, b, pad=37, method=method, tol=1e-11, max_iter=5000, refine_mode=2)
        assert info.status == 0 and info.true_relres <= 1e-11
        assert np.abs(x.numpy() - xe).max() <= 1e-6 * np.abs(xe).max()


def test_config_1024_alpha_error_per_mode(gpu, capsys):
    """1024x1024, one ellipse, 8 % constraints: plain CG, Jacobi-PCG, BF16 and
    FP16 mixed to 1e-6, then FP16 and FP32 with the FP64 master x to 2e-9, and
    the FP64-master solve to 1e-12 that serves as the 4K benchmark's reference;
    max / mean |alpha - alpha_FP64| against the oracle's FP64 Laplacian + FP64
    PCG (1e-12)."""
    p = synth.config_1k()
    x64, it64, rr64 = O.solve_f64(p.image, p.lam, p.target, p.eps, tol=1e-12)
    assert rr64 <= 1e-12
    s = matting.assemble(p.image, p.lam, p.target, p.eps, ASSEMBLY_ACCURATE)
    rows = []
    for name, meth, prec, tol, rm in [("fp32-cg", METHOD_CG, PREC_FP32, 1e-6, 0),
                                      ("fp32-pcg", METHOD_PCG, PREC_FP32, 1e-6, 0),
                                      ("bf16-pcg", METHOD_PCG, PREC_BF16, 1e-6, 0),
                                      ("fp16-pcg", METHOD_PCG, PREC_FP16, 1e-6, 0),
                                      ("fp32-pcg-x64", METHOD_PCG, PREC_FP32, 2e-9, 2),
                                      ("fp16-pcg-x64", METHOD_PCG, PREC_FP16, 2e-9, 2),
                                      # the benchmark's reference alpha, formed the same way at 4K
                                      ("ref-1e-12", METHOD_PCG, PREC_FP32, 1e-12, 1)]:
        x, info = matting.solve(s, method=meth, precision=prec, tol=tol, max_iter=20000, refine_mode=rm)
        d = np.abs(x.numpy().astype(np.float64) - x64)
        rows.append((name, tol, info.iterations, info.true_relres, d.max(), d.mean()))
        assert info.status == 0 and info.true_relres <= tol, name
    with capsys.disabled():
        for r in rows:
            print(f"\n[1024^2] {r[0]:13s} tol {r[1]:.0e}: {r[2]:4d} iterations, true relres {r[3]:.2e}, "
                  f"max |da| {r[4]:.2e}, mean |da| {r[5]:.2e}", end="")
        print()
    for name, tol, its, rr, mx, mean in rows:
        if tol >= 1e-6:
            # at 1e-6 the error is the tolerance's, not the precision's
            assert mx <= 2e-2 and mean <= 5e-5, name
        elif tol >= 1e-9:
            assert mx <= 1e-4 and mean <= 1e-6, name
        else:
            # the FP32 rounding of the matrix moves alpha by far less than 1e-4
            assert mx <= 1e-5, name


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_FP16])
def test_4k_solve_converges(gpu, prec):
    p = synth.config_4k()
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_ACCURATE)
    x, info = matting.solve(s, method=METHOD_PCG, precision=prec, tol=1e-6, max_iter=20000)
    print(f"4K prec={prec}: iterations={info.iterations} true_relres={info.true_relres:.3e}")
    assert info.status == 0 and info.true_relres <= 1e-6
    assert np.isfinite(x.numpy()).all()


def test_autotune_then_solve(gpu, small):
    p, a, b, bands, pad, x_exact = small
    s = _gpu_system(p)
    for prec in (PREC_FP32, PREC_FP16):
        R, ty, pc, us = matting.autotune(s.L, prec, reps=2)
        assert R in ((2, 4) if prec == PREC_FP32 else (4, 8)) and ty in (0, 4, 8, 16) and 1 <= pc <= 8 and us > 0
        x, info = matting.solve(s, method=METHOD_PCG, precision=prec, tol=1e-6)
        assert info.status == 0 and info.true_relres <= 1e-6
        assert np.abs(x.numpy() - x_exact).max() <= 5e-3


@pytest.mark.parametrize("prec,shapes,width", [
    (PREC_FP32, [(2, 0), (4, 0), (4, 4), (4, 8), (4, 16)], 256),
    (PREC_FP16, [(4, 4), (4, 8), (4, 16)], 256),   # symmetric 12-band layout (W % 4 == 0)
    (PREC_FP16, [(4, 0), (8, 0)], 254)])           # full 24-band layout, 1-D tiles
def test_every_tile_shape_same_iteration(gpu, small, prec, shapes, width):
    """Every tile geometry computes the same products: the first iterations of
    a fixed-count CG agree to rounding across shapes."""
    p = small[0] if width == 256 else synth.make_problem(width, 190, 1001, 1, 0.08)
    s = _gpu_system(p)
    outs = []
    for (R, ty) in shapes:
        call("gm_tune_set", p.W, p.N, prec, R, ty, 2)
        x, info = matting.solve(s, method=METHOD_PCG, precision=prec, tol=0, max_iter=30, verify=False)
        outs.append(x.numpy())
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() <= 1e-4
    call("gm_tune_set", p.W, p.N, prec, *shapes[0], 4)


@pytest.mark.parametrize("W,H", [(256, 192), (1000, 37), (1916, 203), (3840, 64), (16640, 8), (512, 3), (8, 64)])
def test_fused_iteration_matches_step_and_update(gpu, W, H):
    """One launch per CG iteration (the update of x and r fused into the next
    step, beta from the expanded <r,r> recurrence, r and Ap alternating between
    two buffers, every other launch deferring its update of x to the next)
    makes the iterates of the two-kernel form: fixed-count runs agree to
    rounding (8 iterations in one chunk; 7 in chunks of 3, so that chunks end
    after deferring and updating launches alike), and runs to a tolerance take
    the same iterations and replacements (every 7 iterations here, so that
    replacements land on odd and even iteration counts) and reach it.  Sizes: ragged right and bottom
    tiles, 15 columns with runs of one tile, 65 columns (more than the
    workgroups per XCD: runs that cross into the next column), an image of 3
    rows (one partial tile per column) and one 8 pixels wide."""
    p = synth.make_problem(W, H, 4321, 1, 0.08)
    s = _gpu_system(p)
    res = {}
    for walk, pc in ((1, 3), (2, 2), (2, 3)):
        call("gm_tune_set_step", W, p.N, PREC_FP16, 4, 4, pc, walk)
        x8, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=0, max_iter=8, verify=False,
                              refine_every=-1)
        x7, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=0, max_iter=7, verify=False,
                              refine_every=-1, check_every=3)
        x, info = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=1e-6, max_iter=5000, refine_every=7)
        res[(walk, pc)] = (x8.numpy(), x7.numpy(), x.numpy(), info)
    call("gm_tune_set_step", W, p.N, PREC_FP16, 4, 4, 2, 2)
    x8r, x7r, xr, ir = res[(1, 3)]
    scale = np.abs(x8r).max()
    assert scale > 0
    for key in ((2, 2), (2, 3)):
        x8, x7, x, info = res[key]
        print(f"\n[{W}x{H} walk {key}] {info.iterations} vs {ir.iterations} iterations, "
              f"{info.replacements} vs {ir.replacements} replacements, true relres {info.true_relres:.3e}", end="")
        assert np.abs(x8 - x8r).max() <= 1e-5 * scale
        assert np.abs(x7 - x7r).max() <= 1e-5 * scale
        assert info.status == 0 and info.true_relres <= 1e-6
        assert abs(info.iterations - ir.iterations) <= 2
        assert np.abs(x - xr).max() <= 1e-3


def test_fused_iteration_restart_and_iteration_cap(gpu, small):
    """The fused form through cgSolve's periodic restart (a residual from
    scratch every 25 iterations, then a fresh first step) and through the
    iteration cap: the same status, iteration count and final residual as the
    two-kernel form."""
    p = small[0]
    s = _gpu_system(p)
    out = {}
    for walk, pc in ((1, 3), (2, 2)):
        call("gm_tune_set_step", p.W, p.N, PREC_FP16, 4, 4, pc, walk)
        x, info = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=1e-6, max_iter=5000, restart=25)
        xc, ic = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=1e-12, max_iter=37)
        out[walk] = (x.numpy(), info, xc.numpy(), ic)
    call("gm_tune_set_step", p.W, p.N, PREC_FP16, 4, 4, 2, 2)
    (xa, ia, xca, ica), (xb, ib, xcb, icb) = out[1], out[2]
    assert ia.status == ib.status == 0 and ib.true_relres <= 1e-6
    assert abs(ia.iterations - ib.iterations) <= 2
    assert np.abs(xa - xb).max() <= 1e-3
    assert ica.status == icb.status == 6 and ica.iterations == icb.iterations == 37  # GM_ERR_NOT_CONVERGED
    assert abs(ica.relres - icb.relres) <= 1e-3 * ica.relres
    assert np.abs(xca - xcb).max() <= 1e-4


@pytest.mark.parametrize("W,H", [(256, 192), (1916, 1023), (2048, 1030), (4096, 300)])
@pytest.mark.parametrize("ty", [4, 8, 16])
def test_column_walk_step_matches_tile_parallel(gpu, W, H, ty):
    """The column-walk form of the symmetric 16-bit step (rows shared with the
    tile above kept in LDS rings) computes the same products as the
    tile-parallel form.  One workgroup per CU, so at the larger sizes every
    workgroup walks several tiles: down one column in a band of rows when each
    XCD has a workgroup per column, else along a run of the column-major order
    that crosses into the next column (4096 wide, 16-row tiles); ragged sizes
    leave partial tiles at the right and bottom edges.  The
    per-workgroup partial sums differ in order between the forms, so the
    iterates agree to rounding, far below what a wrong window row would
    move them."""
    p = synth.make_problem(W, H, 1234, 1, 0.08)
    s = _gpu_system(p)
    outs = []
    for walk in (0, 1):
        call("gm_tune_set_step", W, p.N, PREC_FP16, 4, ty, 1, walk)
        x, info = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=0, max_iter=8, verify=False,
                                refine_every=-1)
        outs.append(x.numpy())
    scale = np.abs(outs[0]).max()
    assert scale > 0
    assert np.abs(outs[1] - outs[0]).max() <= 1e-5 * scale


def _fp16_emulation(p, store=np.float16):
    """Host FP64 emulation of the scaled 16-bit operator: As = S A S with S =
    1/sqrt(a_ii), entries (a_ik s_i) s_j in float, then rounded to `store`
    (FP16 RNE; float32: the fused FP32 iteration's operator) -- for the
    symmetric layout (W % 4 == 0) the upper entry of each pair serves both --
    plus the FP32 matrix A itself (as a sparse matrix), S and b."""
    import scipy.sparse as sp
    W, N = p.W, p.N
    a, b, _ = O.levin_gather_acc_f32(p.image, p.lam, p.target, 1e-7)
    s32 = (np.float32(1) / np.sqrt(a[12][:N].astype(np.float32))).astype(np.float32)
    offs = O.bands25(W)
    rows, cols, vals, vals32 = [], [], [], []
    for k, o in enumerate(offs):
        i = np.arange(N)
        m = (i + o >= 0) & (i + o < N)
        vals32.append((i[m], i[m] + o, a[k][:N][m].astype(np.float64)))
        if k == 12:
            continue
        if W % 4 == 0 and k < 12:
            # lower band: the upper entry of the partner row i + o_k (band 24 - k)
            src = i + o
            v = np.zeros(N, np.float32)
            v[m] = (a[24 - k][src[m]] * s32[src[m]]) * s32[i[m]]
        else:
            j = np.clip(i + o, 0, N - 1)
            v = (a[k][:N] * s32) * s32[j]
        v16 = v.astype(store).astype(np.float64)
        rows.append(i[m]), cols.append(i[m] + o), vals.append(v16[m])
    As = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    As = As + sp.identity(N)
    A = sp.csr_matrix((np.concatenate([v[2] for v in vals32]),
                       (np.concatenate([v[0] for v in vals32]), np.concatenate([v[1] for v in vals32]))), shape=(N, N))
    return As, A, s32.astype(np.float64), b.astype(np.float64)


def _scaled_cg(As, A, sv, b, k, restart=0):
    """k iterations of CG on the scaled system from x = 0 (FP64), sign
    convention r = A x - b, p = -r + beta p; restart: every `restart`
    iterations r from scratch through the FP32 matrix, scaled, and beta = 0."""
    x = np.zeros(b.size)
    r = sv * (A @ x - b)
    pv = np.zeros(b.size)
    rz, beta, since = r @ r, 0.0, 0
    for it in range(k):
        pv = beta * pv - r
        Ap = As @ pv
        al = rz / (pv @ Ap)
        x += al * sv * pv
        r += al * Ap
        rz, beta = r @ r, (r @ r) / rz
        since += 1
        if restart and since >= restart and it + 1 < k:
            r = sv * (A @ x - b)
            rz, beta, since = r @ r, 0.0, 0
    return x


@pytest.mark.parametrize("width", [256, 250])
def test_fp16_operator_matches_emulation(gpu, width):
    """The 16-bit scaled operator the solver applies (the symmetric 12-band
    layout for W % 4 == 0, the full 24-band one otherwise) is the one a host
    emulation builds.  Checked through k fixed-count CG iterations (no residual
    replacement): a wrong lower index, sign or missing band moves x by far
    more than the FP32-vs-FP64 arithmetic difference."""
    p = synth.make_problem(width, 96, 77, 1, 0.08)
    As, A, sv, b = _fp16_emulation(p)
    s = _gpu_system(p)
    for k in (2, 5):
        x, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=0, max_iter=k, verify=False,
                             refine_every=-1)
        xe = _scaled_cg(As, A, sv, b, k)
        assert np.abs(x.numpy() - xe).max() <= 1e-5 * np.abs(xe).max(), k


@pytest.mark.parametrize("walk", [0, 1, 2])
def test_fp16_restart_matches_emulation(gpu, walk):
    """cgSolve's periodic restart on the scaled 16-bit system: the residual is
    formed from scratch as S (A x - b) and the direction restarts, for the
    tile-parallel and column-walk steps and the fused iteration.  The host
    emulation restarts the same way; one iteration after a restart from an
    unscaled residual already moves x by 1e-2 relative."""
    p = synth.make_problem(256, 96, 77, 1, 0.08)
    As, A, sv, b = _fp16_emulation(p)
    s = _gpu_system(p)
    call("gm_tune_set_step", p.W, p.N, PREC_FP16, 4, 4, 2 if walk == 2 else 3, walk)
    for k in (26, 30):
        x, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP16, tol=0, max_iter=k, verify=False,
                             refine_every=-1, restart=25)
        xe = _scaled_cg(As, A, sv, b, k, restart=25)
        assert np.abs(x.numpy() - xe).max() <= 1e-5 * np.abs(xe).max(), k
    call("gm_tune_set_step", p.W, p.N, PREC_FP16, 4, 4, 2, 2)


FUSED32_SIZES = [(256, 192), (1000, 37), (1916, 203), (3840, 64), (16640, 8), (512, 3), (8, 64)]


@pytest.mark.parametrize("W,H", FUSED32_SIZES)
def test_fp32_fused_matches_jacobi_step(gpu, W, H):
    """FP32 Jacobi-PCG as one launch per iteration -- CG on the symmetric
    scaled FP32 operator S A S (12 upper bands, 128-wide column walk) with the
    FP64 residual replacement through A -- makes the iterates of the two-kernel
    Jacobi step + update on A's 25 bands (the same method: x_k of Jacobi-PCG
    and of CG on S A S coincide in exact arithmetic): fixed-count runs agree to
    rounding (8 iterations in one chunk, 7 in chunks of 3: every other launch
    defers its update of x to the next, and chunks end after both kinds), and
    runs to 1e-6 reach it in about as many iterations.  Sizes as
    for the 16-bit fused iteration (W % 4 == 0): ragged tiles, columns with
    runs of one tile, 130 columns, 3 rows, 8 pixels wide."""
    p = synth.make_problem(W, H, 4321, 1, 0.08)
    s = _gpu_system(p)
    res = {}
    for walk, pc in ((0, 3), (2, 2), (2, 3)):
        call("gm_tune_set_step", W, p.N, PREC_FP32, 4, 4, pc, walk)
        x8, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=0, max_iter=8, verify=False,
                              refine_every=-1)
        x7, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=0, max_iter=7, verify=False,
                              refine_every=-1, check_every=3)
        x, info = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=1e-6, max_iter=5000, refine_every=7)
        res[(walk, pc)] = (x8.numpy(), x7.numpy(), x.numpy(), info)
    call("gm_tune_set_step", W, p.N, PREC_FP32, 4, 4, 2, 2)
    x8r, x7r, xr, ir = res[(0, 3)]
    scale = np.abs(x8r).max()
    assert scale > 0 and ir.status == 0
    for key in ((2, 2), (2, 3)):
        x8, x7, x, info = res[key]
        print(f"\n[{W}x{H} FP32 fused {key}] {info.iterations} vs {ir.iterations} iterations, "
              f"{info.replacements} replacements, true relres {info.true_relres:.3e}", end="")
        assert np.abs(x8 - x8r).max() <= 1e-5 * scale
        assert np.abs(x7 - x7r).max() <= 1e-5 * scale
        assert info.status == 0 and info.true_relres <= 1e-6 and info.replacements >= 1
        assert abs(info.iterations - ir.iterations) <= max(3, 0.05 * ir.iterations)
        assert np.abs(x - xr).max() <= 1e-3


@pytest.mark.parametrize("W,H", [(256, 96), (1000, 37), (8, 64)])
def test_fp32_fused_operator_matches_emulation(gpu, W, H):
    """The fused FP32 iteration's operator is the host emulation's symmetric
    scaled FP32 one (upper entry of each pair for both, unit diagonal): k
    fixed-count iterations without replacement against FP64 CG on it.  A
    wrong lower index, sign or missing band moves x far more than 1e-5."""
    p = synth.make_problem(W, H, 77, 1, 0.08)
    As, A, sv, b = _fp16_emulation(p, np.float32)
    s = _gpu_system(p)
    call("gm_tune_set_step", W, p.N, PREC_FP32, 4, 4, 2, 2)
    for k in (2, 5, 10):
        x, _ = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=0, max_iter=k, verify=False,
                             refine_every=-1)
        xe = _scaled_cg(As, A, sv, b, k)
        assert np.abs(x.numpy() - xe).max() <= 1e-5 * np.abs(xe).max(), k


def test_fp32_fused_restart_and_iteration_cap(gpu, small):
    """The fused FP32 form through cgSolve's periodic restart and through the
    iteration cap: the status, iteration count and final residual of the
    Jacobi step + update."""
    p = small[0]
    s = _gpu_system(p)
    out = {}
    for walk, pc in ((0, 3), (2, 2)):
        call("gm_tune_set_step", p.W, p.N, PREC_FP32, 4, 4, pc, walk)
        x, info = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=1e-6, max_iter=5000, restart=25)
        xc, ic = matting.solve(s, method=METHOD_PCG, precision=PREC_FP32, tol=1e-12, max_iter=37)
        out[walk] = (x.numpy(), info, xc.numpy(), ic)
    call("gm_tune_set_step", p.W, p.N, PREC_FP32, 4, 4, 2, 2)
    (xa, ia, xca, ica), (xb, ib, xcb, icb) = out[0], out[2]
    assert ia.status == ib.status == 0 and ib.true_relres <= 1e-6
    assert abs(ia.iterations - ib.iterations) <= 3
    assert np.abs(xa - xb).max() <= 1e-3
    assert ica.status == icb.status == 6 and ica.iterations == icb.iterations == 37  # GM_ERR_NOT_CONVERGED
    assert abs(ica.relres - icb.relres) <= 1e-3 * ica.relres
    assert np.abs(xca - xcb).max() <= 1e-4
