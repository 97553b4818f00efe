"""ctypes front end of the CPU oracle (oracle/oracle.c) -- TEST INFRASTRUCTURE.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this.
Matrices are numpy arrays of shape (nbands, apitch) (diagonal-major, the
reference's layout, BandedMatrix.h:25-32); vectors that a banded product reads
are padded numpy arrays of length N + 2*pad whose element `pad` is x[0].
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
vp = C.c_void_p
sz = C.c_size_t
ci = C.c_int
cf = C.c_float
cd = C.c_double


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(C.CDLL(path))
    return _LIB


def baseline_lib(path):
    """The CPU-baseline build of oracle.c (oracle/Makefile target baseline:
    -O3 -march=native, FP contraction allowed) at `path`."""
    return _bind(C.CDLL(path))


def _bind(L):
    L.orc_bands25.argtypes = [ci, i32p]
    L.orc_pad.argtypes = [ci]
    L.orc_pad.restype = ci
    L.orc_levin_scatter_f32.argtypes = [f32p, sz, f32p, f32p, ci, sz, vp, vp, cf, ci, ci]
    L.orc_levin_scatter_u8.argtypes = [f32p, sz, f32p, f32p, ci, sz, u8p, cf, cf, ci, ci]
    L.orc_levin_gather_f32.argtypes = [f32p, sz, f32p, vp, f32p, ci, sz, vp, vp, cf, ci, ci]
    L.orc_levin_gather_acc_f32.argtypes = [f32p, sz, f32p, vp, f32p, ci, sz, vp, vp, cf, ci, ci]
    L.orc_levin_f64.argtypes = [f64p, sz, f64p, f32p, ci, sz, vp, vp, cd, ci, ci]
    L.orc_levin_dense_f64.argtypes = [f64p, f32p, ci, sz, cd, ci, ci]
    for nm in ("orc_bm_ax_f32",):
        getattr(L, nm).argtypes = [vp, vp, sz, i32p, ci, ci, vp]
    L.orc_bm_axpy_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, vp, ci]
    L.orc_bm_ax_f64.argtypes = [vp, vp, sz, i32p, ci, ci, vp]
    L.orc_dot_f32.argtypes = [f32p, f32p, ci]
    L.orc_dot_f32.restype = cd
    L.orc_cg_ref_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, ci, ci, ci]
    L.orc_gd_ref_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, ci, ci]
    L.orc_pcg_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, ci, cd, ci, ci, ci, C.POINTER(cd)]
    L.orc_pcg_f32.restype = ci
    L.orc_pcg_f64.argtypes = [vp, vp, sz, i32p, ci, ci, vp, ci, cd, ci, C.POINTER(cd)]
    L.orc_pcg_f64.restype = ci
    L.orc_true_relres_f32.argtypes = [vp, sz, i32p, ci, ci, vp, vp]
    L.orc_true_relres_f32.restype = cd
    L.orc_resid_data_f64.argtypes = [f64p, vp, sz, i32p, ci, ci, vp, vp, vp]
    L.orc_gs_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, sz]
    L.orc_jacobi_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, sz, ci, cf]
    L.orc_mse_clamped.argtypes = [f32p, f32p, ci, ci]
    L.orc_mse_clamped.restype = cd
    L.orc_num_threads.restype = ci
    L.orc_set_num_threads.argtypes = [ci]
    L.orc_levin_backward_f64.argtypes = [f64p, f32p, ci, sz, f64p, f64p, cd, ci, ci]
    L.orc_baseline_pcg_f32.argtypes = [vp, vp, sz, i32p, ci, ci, vp, ci, ci, cd, C.POINTER(ci)]
    L.orc_baseline_pcg_f32.restype = cd
    return L


def _ptr(a: np.ndarray, off: int = 0) -> int:
    return a.ctypes.data + off * a.itemsize


def bands25(W: int) -> np.ndarray:
    b = np.zeros(25, np.int32)
    lib().orc_bands25(W, b)
    return b


def pad_for(W: int) -> int:
    return lib().orc_pad(W)


def padded(x: np.ndarray, pad: int) -> np.ndarray:
    out = np.zeros(x.size + 2 * pad, x.dtype)
    out[pad:pad + x.size] = x.ravel()
    return out


def _img(image: np.ndarray):
    im = np.ascontiguousarray(image, np.float32)
    H, W, nch = im.shape
    return im, W, H, nch


def _opt(a, dt):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dt).ravel()
    return arr, arr.ctypes.data


def levin_scatter_f32(image, lam=None, tgt=None, eps=1e-7, apitch=None):
    """Reference-semantics FP32 Laplacian (Matting.cu:39-219): returns (a, b)."""
    im, W, H, nch = _img(image)
    N = W * H
    ap = apitch or N
    a = np.zeros((25, ap), np.float32)
    b = np.zeros(N, np.float32)
    la, lp = _opt(lam, np.float32)
    ta, tp = _opt(tgt, np.float32)
    lib().orc_levin_scatter_f32(a, ap, b, im, nch, W, lp, tp, eps, W, H)
    return a, b


def levin_scatter_u8(image, scribs, eps=1e-2, gamma=1.0):
    im, W, H, nch = _img(image)
    N = W * H
    a = np.zeros((25, N), np.float32)
    b = np.zeros(N, np.float32)
    lib().orc_levin_scatter_u8(a, N, b, im, nch, W, np.ascontiguousarray(scribs, np.uint8), eps,
                               gamma, W, H)
    return a, b


def levin_gather_f32(image, lam=None, tgt=None, eps=1e-7, apitch=None):
    """Same numbers as levin_scatter_f32, row-parallel; returns (a, b, dinv)."""
    im, W, H, nch = _img(image)
    N = W * H
    ap = apitch or N
    a = np.zeros((25, ap), np.float32)
    b = np.zeros(N, np.float32)
    dinv = np.zeros(N, np.float32)
    la, lp = _opt(lam, np.float32)
    ta, tp = _opt(tgt, np.float32)
    lib().orc_levin_gather_f32(a, ap, b, dinv.ctypes.data, im, nch, W, lp, tp, eps, W, H)
    return a, b, dinv


def levin_gather_acc_f32(image, lam=None, tgt=None, eps=1e-7, apitch=None):
    """Accurate-rounding FP32 assembly (FP64 window statistics, FMA-contracted
    whitening, float centring on the head of the window mean); returns
    (a, b, dinv).  The GPU's GM_ASSEMBLY_ACCURATE operation for operation."""
    im, W, H, nch = _img(image)
    N = W * H
    ap = apitch or N
    a = np.zeros((25, ap), np.float32)
    b = np.zeros(N, np.float32)
    dinv = np.zeros(N, np.float32)
    la, lp = _opt(lam, np.float32)
    ta, tp = _opt(tgt, np.float32)
    lib().orc_levin_gather_acc_f32(a, ap, b, dinv.ctypes.data, im, nch, W, lp, tp, eps, W, H)
    return a, b, dinv


def levin_f64(image, lam=None, tgt=None, eps=1e-7):
    im, W, H, nch = _img(image)
    N = W * H
    a = np.zeros((25, N), np.float64)
    b = np.zeros(N, np.float64)
    la, lp = _opt(lam, np.float64)
    ta, tp = _opt(tgt, np.float64)
    lib().orc_levin_f64(a, N, b, im, nch, W, lp, tp, eps, W, H)
    return a, b


def levin_dense_f64(image, eps=1e-7):
    im, W, H, nch = _img(image)
    N = W * H
    L = np.zeros((N, N), np.float64)
    lib().orc_levin_dense_f64(L, im, nch, W, eps, W, H)
    return L


def bm_ax_f32(a, bands, xpad, pad):
    rows = a.shape[1] if xpad.size - 2 * pad > a.shape[1] else xpad.size - 2 * pad
    y = np.zeros(rows, np.float32)
    lib().orc_bm_ax_f32(_ptr(y), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                        len(bands), rows, _ptr(xpad, pad))
    return y


def bm_ax_f64(a, bands, xpad, pad):
    rows = xpad.size - 2 * pad
    y = np.zeros(rows, np.float64)
    lib().orc_bm_ax_f64(_ptr(y), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                        len(bands), rows, _ptr(xpad, pad))
    return y


def dot_f32(x, y) -> float:
    return lib().orc_dot_f32(np.ascontiguousarray(x, np.float32), np.ascontiguousarray(y, np.float32),
                             x.size)


def cg_ref_f32(a, bands, b, x0, pad, iterations, restart=101):
    """main.cu:458-526 restated; returns alpha (unpadded)."""
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    lib().orc_cg_ref_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                         len(bands), N, _ptr(b), pad, iterations, restart)
    return xp[pad:pad + N].copy()


def gd_ref_f32(a, bands, b, x0, pad, iterations):
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    lib().orc_gd_ref_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                         len(bands), N, _ptr(b), pad, iterations)
    return xp[pad:pad + N].copy()


def pcg_f32(a, bands, b, x0, pad, tol=1e-6, maxit=10000, restart=0, precond=True):
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    rr = cd(0)
    it = lib().orc_pcg_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                           len(bands), N, _ptr(b), pad, tol, maxit, restart, int(precond), C.byref(rr))
    return xp[pad:pad + N].copy(), it, rr.value


def pcg_f64(a64, bands, b64, x0, pad, tol=1e-12, maxit=100000):
    N = b64.size
    xp = padded(np.asarray(x0, np.float64), pad)
    rr = cd(0)
    it = lib().orc_pcg_f64(_ptr(xp, pad), _ptr(a64), a64.shape[1], np.ascontiguousarray(bands, np.int32),
                           len(bands), N, _ptr(b64), pad, tol, maxit, C.byref(rr))
    return xp[pad:pad + N].copy(), it, rr.value


def true_relres_f32(a, bands, b, x, pad):
    xp = padded(np.asarray(x, np.float32), pad)
    b = np.ascontiguousarray(b, np.float32)
    return lib().orc_true_relres_f32(_ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                                     len(bands), b.size, _ptr(b), _ptr(xp, pad))


def resid_data_f64(a, bands, lam, tgt, x, pad):
    """(||A x - b||, ||b||) in FP64 for the FP32 system with constraint weights
    lam and targets tgt on a Laplacian a without data term (a_ii + lam_i and
    lam_i * tgt_i rounded to float, as a BandedMatrix stores them)."""
    xp = padded(np.asarray(x, np.float32).ravel(), pad)
    la = np.ascontiguousarray(lam, np.float32).ravel()
    tg = np.ascontiguousarray(tgt, np.float32).ravel()
    out = np.zeros(2, np.float64)
    lib().orc_resid_data_f64(out, _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32), len(bands), la.size,
                             _ptr(la), _ptr(tg), _ptr(xp, pad))
    return float(np.sqrt(out[0])), float(np.sqrt(out[1]))


def gs_f32(a, bands, b, x0, pad, iterations):
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    lib().orc_gs_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                     len(bands), N, _ptr(b), iterations)
    return xp[pad:pad + N].copy()


def jacobi_f32(a, bands, b, x0, pad, iterations, omega=2.0 / 3.0):
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    lib().orc_jacobi_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                         len(bands), N, _ptr(b), iterations, pad, omega)
    return xp[pad:pad + N].copy()


def mse_clamped(alpha, gt):
    H, W = gt.shape
    return lib().orc_mse_clamped(np.ascontiguousarray(alpha, np.float32).ravel(),
                                 np.ascontiguousarray(gt, np.float32).ravel(), W, H)


def baseline_pcg_f32(a, bands, b, x0, pad, maxit, tol=0.0, L=None):
    """OpenMP Jacobi-PCG (the CPU baseline) until ||r||/||b|| <= tol or maxit
    iterations (tol 0: exactly maxit): (x, iterations, relres).  L: the
    library to run it in (default the oracle's)."""
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    it = C.c_int(0)
    rr = (L or lib()).orc_baseline_pcg_f32(_ptr(xp, pad), _ptr(a), a.shape[1], np.ascontiguousarray(bands, np.int32),
                                           len(bands), N, _ptr(b), pad, maxit, tol, C.byref(it))
    return xp[pad:pad + N].copy(), it.value, rr


def levin_backward_f64(image, mu, alpha, eps=1e-7):
    """d/d image of sum_{i!=j} mu_i (alpha_i - alpha_j) L_ij (the assembly
    adjoint); returns (H, W, 3) float64."""
    im, W, H, nch = _img(image)
    g = np.zeros(W * H * 3, np.float64)
    lib().orc_levin_backward_f64(g, im, nch, W, np.ascontiguousarray(mu, np.float64).ravel(),
                                 np.ascontiguousarray(alpha, np.float64).ravel(), eps, W, H)
    return g.reshape(H, W, 3)


def num_threads() -> int:
    return lib().orc_num_threads()


def solve_f64(image, lam, tgt, eps=1e-7, tol=1e-12, maxit=200000):
    """The FP64 accuracy oracle: FP64 Laplacian + FP64 Jacobi-PCG.  Returns alpha (N,)."""
    H, W = image.shape[:2]
    a64, b64 = levin_f64(image, lam, tgt, eps)
    bands = bands25(W)
    pad = pad_for(W)
    x, it, rr = pcg_f64(a64, bands, b64, np.zeros(W * H), pad, tol, maxit)
    return x, it, rr


# --------------------------------------------------------------------------
# The reference's own host solvers, compiled from /root/reference/Solve.cpp
# into oracle/_ref (see oracle/Makefile).  C++ symbols are looked up by name.

class _RefBM(C.Structure):
    # BandedMatrix.h:23-38
    _fields_ = [("a", C.c_void_p), ("apitch", C.c_size_t), ("bands", C.c_void_p),
                ("nbands", C.c_int), ("rows", C.c_int), ("cols", C.c_int)]


def ref_lib():
    """The reference Solve.cpp build, or None when it was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_solve.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path)
    return _REF


def _ref_sym(name):
    import re
    L = ref_lib()
    out = subprocess.run(["nm", "-D", os.path.join(_HERE, "_ref", "libref_solve.so")],
                         capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        m = re.search(r"\s(_Z\d+" + name + r"\S*)$", line)
        if m:
            return getattr(L, m.group(1))
    raise KeyError(name)


def ref_gauss_seidel(a, bands, b, x0, pad, iterations):
    """Solve.cpp:3-33 (reference code) on a padded copy of x0."""
    N = b.size
    xp = padded(np.asarray(x0, np.float32), pad)
    bd = np.ascontiguousarray(bands, np.int32)
    M = _RefBM(_ptr(a), a.shape[1], bd.ctypes.data, len(bands), N, N)
    f = _ref_sym("gaussSeidel_host")
    f.argtypes = [C.c_void_p, C.POINTER(_RefBM), C.c_void_p, C.c_size_t]
    f(_ptr(xp, pad), C.byref(M), _ptr(np.ascontiguousarray(b, np.float32)), iterations)
    return xp[pad:pad + N].copy()


def ref_jacobi(a, bands, b, x0, pad, iterations, omega=2.0 / 3.0):
    """Solve.cpp:35-85 (reference code).  Only even iteration counts: after an odd
    count the reference frees the caller's buffer (Solve.cpp:53-63)."""
    if iterations % 2:
        raise ValueError("reference jacobi_host is only safe for even iteration counts")
    N = b.size
    # the reference deletes its own buffer with delete[]; the caller's buffer
    # must stay ours, which holds for even counts
    xp = padded(np.asarray(x0, np.float32), pad)
    bd = np.ascontiguousarray(bands, np.int32)
    M = _RefBM(_ptr(a), a.shape[1], bd.ctypes.data, len(bands), N, N)
    f = _ref_sym("jacobi_host")
    f.argtypes = [C.c_void_p, C.POINTER(_RefBM), C.c_void_p, C.c_size_t, C.c_int, C.c_float]
    f(_ptr(xp, pad), C.byref(M), _ptr(np.ascontiguousarray(b, np.float32)), iterations, pad, omega)
    return xp[pad:pad + N].copy()


def matte_loss_grad_f64(image, lam, tgt, alpha_gt, eps=1e-7, tol=1e-14):
    """FP64 host reference of loss = ||alpha - alpha_gt||^2 and its gradients
    with respect to the image, lam and tgt (forward solve, adjoint solve,
    assembly adjoint).  Returns (loss, alpha, d_image (H,W,3), d_lam, d_tgt)."""
    H, W = image.shape[:2]
    a64, b64 = levin_f64(image, lam, tgt, eps)
    bands = bands25(W)
    pad = pad_for(W)
    al, _, _ = pcg_f64(a64, bands, b64, np.zeros(W * H), pad, tol, 1000000)
    r = al - np.asarray(alpha_gt, np.float64).ravel()
    loss = float(r @ r)
    mu, _, _ = pcg_f64(a64, bands, 2 * r, np.zeros(W * H), pad, tol, 1000000)
    gim = levin_backward_f64(image, mu, al, eps)
    lam64 = np.asarray(lam, np.float64).ravel()
    tg64 = np.asarray(tgt, np.float64).ravel()
    return loss, al, gim, mu * (tg64 - al), mu * lam64


def matte_loss_f64(image, lam, tgt, alpha_gt, eps=1e-7, tol=1e-14):
    H, W = image.shape[:2]
    a64, b64 = levin_f64(image, lam, tgt, eps)
    al, _, _ = pcg_f64(a64, bands25(W), b64, np.zeros(W * H), pad_for(W), tol, 1000000)
    r = al - np.asarray(alpha_gt, np.float64).ravel()
    return float(r @ r)


def matte_loss_fd_f64(image, lam, tgt, alpha_gt, coords, h=1e-4, eps=1e-7, tol=1e-11, workers=8):
    """FP64 central finite differences of loss = ||alpha - alpha_gt||^2, the
    loss of matte_loss_grad_f64, at `coords`: ("rgb", y, x, c) for an image
    channel, ("lam", i) for a constraint weight.  For each coordinate the
    parameter moves to theta +- h (the image is float32, so the step is the
    float32 difference actually taken).  The change of the solution, delta =
    alpha(theta +- h) - alpha, is solved for directly, A(theta +- h) delta =
    b(theta +- h) - A(theta +- h) alpha, and the loss change is 2 (alpha -
    alpha_gt) . delta + delta . delta: the same central difference as two full
    solves and a subtraction of losses, without the cancellation.  Returns the
    derivatives in the order of coords."""
    import concurrent.futures as cf
    image = np.ascontiguousarray(image, np.float32)
    H, W = image.shape[:2]
    N = W * H
    bands = bands25(W)
    pad = pad_for(W)
    lam = np.ascontiguousarray(lam, np.float32).ravel()
    tgt = np.ascontiguousarray(tgt, np.float32).ravel()
    a64, b64 = levin_f64(image, lam, tgt, eps)
    al, _, _ = pcg_f64(a64, bands, b64, np.zeros(N), pad, 1e-14, 1000000)
    r = al - np.asarray(alpha_gt, np.float64).ravel()
    alp = padded(al, pad)

    def dloss(img, la):
        a, b = levin_f64(img, la, tgt, eps)
        rhs = np.zeros(N)
        lib().orc_bm_ax_f64(_ptr(rhs), _ptr(a), N, np.ascontiguousarray(bands, np.int32), 25, N, _ptr(alp, pad))
        rhs = b - rhs
        d, _, _ = pcg_f64(a, bands, rhs, np.zeros(N), pad, tol, 1000000)
        return 2.0 * float(r @ d) + float(d @ d)

    def one(cd):
        lib().orc_set_num_threads(1)
        if cd[0] == "rgb":
            _, y, x, c = cd
            up, dn = image.copy(), image.copy()
            up[y, x, c] = np.float32(float(image[y, x, c]) + h)
            dn[y, x, c] = np.float32(float(image[y, x, c]) - h)
            return (dloss(up, lam) - dloss(dn, lam)) / (float(up[y, x, c]) - float(dn[y, x, c]))
        _, i = cd
        up, dn = lam.copy(), lam.copy()
        up[i] = np.float32(float(lam[i]) + h)
        dn[i] = np.float32(float(lam[i]) - h)
        return (dloss(image, up) - dloss(image, dn)) / (float(up[i]) - float(dn[i]))

    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        out = list(ex.map(one, coords))
    lib().orc_set_num_threads(num_threads())
    return np.array(out)
