"""GPU parity of the building blocks through the C ABI: banded SpMV (generic
and LDS-tiled Levin kernels), BLAS-1, and the Levin assembly (bit-identical to
the oracle's restatement of Matting.cu:39-219)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import torch

from gpumatting_amd import ASSEMBLY_ACCURATE, ASSEMBLY_REFERENCE, _lib, matting, synth
from gpumatting_amd._lib import BandedMatrix, call, lib
from gpumatting_amd.matting import DeviceMatrix, Padded
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
S0 = C.c_void_p(0)


def dptr(t):
    return C.c_void_p(t.data_ptr())


def test_cpp_bandedtest_and_vectortest(gpu):
    for prog in ("bandedtest", "vectortest", "sparsetest"):
        exe = os.path.join(ROOT, "tests", "cpp", "bin", prog)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0 and "Test PASSED" in r.stdout, r.stdout + r.stderr


def test_bandedtest_known_answer_generic_path(gpu):
    N = 1_000_000
    bands = np.array([-10, 0, 10], np.int32)
    a = np.zeros((3, N), np.float32)
    i = np.arange(N)
    a[0, i >= 10] = 1
    a[1] = 1
    a[2, i + 10 < N] = 1
    M = DeviceMatrix(N, bands, a)
    x = Padded(N, 10, init=np.arange(N, dtype=np.float32))
    y = matting.bm_ax(M, x).cpu().numpy().astype(np.float64)
    exp = np.where(i < 10, 10.0 + 2 * i, np.where(i < N - 10, 3.0 * i, 1999970.0 + 2 * (i - (N - 10))))
    assert np.array_equal(y, exp)


def test_vectortest_known_answer_and_determinism(gpu):
    n = 1_000_000
    x = torch.ones(n, device="cuda")
    y = torch.full((n,), 2.0, device="cuda")
    r = torch.zeros(1, device="cuda")
    call("gm_inner_prod", dptr(r), dptr(x), dptr(y), n, S0)
    assert r.item() == 2e6
    g = torch.Generator(device="cpu").manual_seed(0)
    u = torch.randn(3_000_017, generator=g).cuda()
    v = torch.randn(3_000_017, generator=g).cuda()
    outs = []
    for _ in range(3):
        call("gm_inner_prod", dptr(r), dptr(u), dptr(v), u.numel(), S0)
        outs.append(r.item())
    assert outs[0] == outs[1] == outs[2]
    exact = float(np.dot(u.cpu().numpy().astype(np.float64), v.cpu().numpy().astype(np.float64)))
    assert abs(outs[0] - exact) <= 1e-4 * abs(exact) + 1e-2


def test_vector_ops(gpu):
    n = 100_003
    x = torch.randn(n, device="cuda")
    y = torch.randn(n, device="cuda")
    r = torch.empty(n, device="cuda")
    call("gm_vec_add", dptr(r), dptr(x), dptr(y), n, S0)
    assert torch.equal(r, x + y)
    call("gm_vec_sub", dptr(r), dptr(x), dptr(y), n, S0)
    assert torch.equal(r, x - y)
    call("gm_vec_scale_const", dptr(r), dptr(x), C.c_float(-1.5), n, S0)
    assert torch.equal(r, x * -1.5)
    s = torch.tensor([0.25], device="cuda")
    call("gm_vec_scale", dptr(r), dptr(x), dptr(s), n, S0)
    assert torch.equal(r, x * 0.25)


LEVIN_SIZES = [(1, 1), (2, 2), (3, 3), (4, 3), (5, 4), (17, 13), (64, 4), (65, 5), (67, 35), (256, 192)]


@pytest.mark.parametrize("W,H", LEVIN_SIZES)
def test_levin_reference_bitwise(gpu, W, H):
    """GM_ASSEMBLY_REFERENCE == the scatter restatement of Matting.cu, bit for bit."""
    p = synth.make_problem(W, H, 100 + W + H, 1, 0.08)
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_REFERENCE)
    a_ref, b_ref = O.levin_scatter_f32(p.image, p.lam, p.target, 1e-7)
    a = s.L.to_numpy()
    assert np.array_equal(a.view(np.uint32), a_ref.view(np.uint32))
    assert np.array_equal(s.b.cpu().numpy().view(np.uint32), b_ref.view(np.uint32))
    d = a_ref[12]
    dinv = np.where(d != 0, np.float32(1) / np.where(d != 0, d, 1), 0).astype(np.float32)
    assert np.array_equal(s.dinv.cpu().numpy(), dinv)


@pytest.mark.parametrize("W,H", [(3, 3), (17, 13), (67, 35), (256, 192)])
def test_levin_accurate_bitwise(gpu, W, H):
    p = synth.make_problem(W, H, 200 + W, 1, 0.08)
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_ACCURATE)
    a_ref, b_ref, d_ref = O.levin_gather_acc_f32(p.image, p.lam, p.target, 1e-7)
    assert np.array_equal(s.L.to_numpy().view(np.uint32), a_ref.view(np.uint32))
    assert np.array_equal(s.b.cpu().numpy(), b_ref)
    assert np.array_equal(s.dinv.cpu().numpy(), d_ref)


def test_levin_rgba_pitch_and_u8_rule_on_reference_data(gpu):
    """The reference's own call on a crop of data/05 (eps 1e-2, gamma 1, float4
    RGBA input, u8 scribbles) -- GPU u8 path and the hostLevinLaplacian drop-in."""
    z = np.load(os.path.join(ROOT, "tests", "golden", "data05_crop.npz"))
    f = z["rgb"].astype(np.float32) / np.float32(int(z["maxval"]))
    im4 = np.ascontiguousarray(np.concatenate([f, np.ones(f.shape[:2] + (1,), np.float32)], axis=2))
    sc = np.ascontiguousarray(z["scribs"])
    H, W = sc.shape
    a_ref, b_ref = O.levin_scatter_u8(im4, sc, eps=1e-2, gamma=1.0)
    s = matting.assemble_u8(im4, sc, eps=1e-2, gamma=1.0)
    assert np.array_equal(s.L.to_numpy().view(np.uint32), a_ref.view(np.uint32))
    assert np.array_equal(s.b.cpu().numpy(), b_ref)
    # host drop-in: host L with apitch > rows, image with pitch > W
    pitch = W + 5
    imp = np.zeros((H, pitch, 4), np.float32)
    imp[:, :W] = im4
    scp = np.full((H, pitch), 7, np.uint8)
    scp[:, :W] = sc
    N = W * H
    ap = N + 13
    a = np.zeros((25, ap), np.float32)
    bands = O.bands25(W)
    hL = BandedMatrix(a.ctypes.data, ap, bands.ctypes.data, 25, N, N)
    b = np.zeros(N, np.float32)
    call("gm_host_levin_laplacian", C.byref(hL), b.ctypes.data, C.c_float(1e-2), imp.ctypes.data, scp.ctypes.data,
         W, H, pitch)
    assert np.array_equal(a[:, :N].view(np.uint32), a_ref.view(np.uint32))
    assert np.array_equal(b, b_ref)


@pytest.mark.parametrize("W,H", [(5, 4), (67, 35), (256, 192), (1000, 7)])
def test_tiled_spmv_vs_oracle(gpu, W, H):
    p = synth.make_problem(W, H, 300 + W, 1, 0.08)
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_ACCURATE)
    a = s.L.to_numpy()
    rng = np.random.default_rng(W)
    x = rng.standard_normal(p.N).astype(np.float32)
    y = rng.standard_normal(p.N).astype(np.float32)
    xp = Padded(p.N, s.pad, init=x)
    pad = s.pad
    ref = O.bm_ax_f64(a.astype(np.float64), O.bands25(W), O.padded(x.astype(np.float64), pad), pad)
    scale = np.abs(a).sum(0) * np.abs(x).max() + 1e-30
    got = matting.bm_ax(s.L, xp).cpu().numpy()
    assert (np.abs(got - ref) / scale).max() <= 1e-6
    yt = torch.from_numpy(y).cuda()
    got = matting.bm_ax(s.L, xp, yt, add=True).cpu().numpy()
    assert (np.abs(got - (ref + y)) / (scale + np.abs(y))).max() <= 1e-6
    got = matting.bm_ax(s.L, xp, yt, add=False).cpu().numpy()
    assert (np.abs(got - (ref - y)) / (scale + np.abs(y))).max() <= 1e-6


SPMV_FORMS = [(2, 0, 0), (4, 0, 0), (4, 4, 0), (4, 8, 0), (4, 16, 0), (4, 4, 1), (4, 8, 1), (2, 8, 1), (2, 16, 1),
              (8, 4, 2), (12, 4, 2), (16, 4, 2)]


@pytest.mark.parametrize("W,H", [(64, 20), (256, 192), (96, 7), (1000, 37)])
def test_every_spmv_form_same_product(gpu, W, H):
    """Every form of the plain product (1-D and 2-D staged tiles, the LDS-free
    direct tiles with their guarded edge tiles, the long-run tiles) sums the 25 products of a row in
    the same order: bit-identical results, and the FP64 oracle's to 1e-6."""
    p = synth.make_problem(W, H, 400 + W, 1, 0.08)
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_ACCURATE)
    a = s.L.to_numpy()
    rng = np.random.default_rng(H)
    x = rng.standard_normal(p.N).astype(np.float32)
    xp = Padded(p.N, s.pad, init=x)
    ref = O.bm_ax_f64(a.astype(np.float64), O.bands25(W), O.padded(x.astype(np.float64), s.pad), s.pad)
    scale = np.abs(a).sum(0) * np.abs(x).max() + 1e-30
    outs = []
    for (R, ty, direct) in SPMV_FORMS:
        if W % (4 if direct == 2 else R):
            continue
        call("gm_spmv_tune_set", W, p.N, R, ty, direct, 4)
        got = matting.bm_ax(s.L, xp).cpu().numpy()
        assert (np.abs(got - ref) / scale).max() <= 1e-6, (R, ty, direct)
        outs.append(got)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.uint32), outs[0].view(np.uint32))
    call("gm_spmv_tune_set", W, p.N, 4, 4, 0, 8)


def test_spmv_properties_at_4k(gpu):
    """Size-independent properties at 3840x2160: linearity, symmetry
    <x, Ay> = <Ax, y>, and L 1 = lambda D 1 (zero row sums of L)."""
    p = synth.config_4k()
    s = matting.assemble(p.image, p.lam, p.target, 1e-7, ASSEMBLY_ACCURATE)
    N = p.N
    g = torch.Generator(device="cuda").manual_seed(1)
    x = Padded(N, s.pad, init=torch.rand(N, device="cuda", generator=g))
    y = Padded(N, s.pad, init=torch.rand(N, device="cuda", generator=g))
    xy = Padded(N, s.pad, init=x.t + y.t)
    Ax, Ay, Axy = matting.bm_ax(s.L, x), matting.bm_ax(s.L, y), matting.bm_ax(s.L, xy)
    lin = (Axy - Ax - Ay).abs().max().item()
    assert lin <= 1e-4 * max(Ax.abs().max().item(), 1.0)
    lhs = torch.dot(x.t.double(), Ay.double()).item()
    rhs = torch.dot(Ax.double(), y.t.double()).item()
    assert abs(lhs - rhs) <= 1e-6 * (abs(lhs) + abs(rhs))
    one = Padded(N, s.pad, init=torch.ones(N, device="cuda"))
    r = matting.bm_ax(s.L, one)
    lam = torch.from_numpy(p.lam.ravel()).cuda()
    assert (r - lam).abs().max().item() <= 1e-4
