"""Device-level Python API over the C ABI, with PyTorch tensors as device
memory (allocation and streams only -- every computation is a library call).

Vectors that a banded product reads at i + offset are "padded": a zeroed
tensor of N + 2*pad floats whose element `pad` is x[0] (main.cu:228-229,
Vector.cu:23-30).  `Padded.ptr` is the address of x[0].
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import BandedMatrix, SolveInfo, SolveOpts, call, lib


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t: torch.Tensor | None):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def levin_bands(W: int) -> np.ndarray:
    b = np.zeros(25, np.int32)
    lib().gm_levin_bands(W, b.ctypes.data)
    return b


def levin_pad(W: int) -> int:
    return lib().gm_levin_pad(W)


class Padded:
    """A zero-padded FP32 device vector."""

    def __init__(self, n: int, pad: int, device="cuda", init: torch.Tensor | np.ndarray | None = None):
        self.n, self.pad = n, pad
        self.buf = torch.zeros(n + 2 * pad, dtype=torch.float32, device=device)
        if init is not None:
            src = torch.as_tensor(np.asarray(init, np.float32).ravel() if isinstance(init, np.ndarray) else init)
            self.buf[pad:pad + n] = src.reshape(-1).to(device=device, dtype=torch.float32)

    @property
    def t(self) -> torch.Tensor:
        return self.buf[self.pad:self.pad + self.n]

    @property
    def ptr(self):
        return C.c_void_p(self.buf.data_ptr() + 4 * self.pad)

    def numpy(self) -> np.ndarray:
        return self.t.detach().cpu().numpy()


class DeviceMatrix:
    """A device BandedMatrix allocated by the library (pitch rounded to 64,
    zero padding).  Free with .free() (or let the object go)."""

    def __init__(self, rows: int, bands, host_a: np.ndarray | None = None):
        self.m = BandedMatrix()
        bands = np.ascontiguousarray(bands, np.int32)
        if host_a is None:
            call("gm_bm_device_alloc", C.byref(self.m), rows, len(bands), bands.ctypes.data)
        else:
            a = np.ascontiguousarray(host_a, np.float32)
            h = BandedMatrix(a.ctypes.data, a.shape[1], bands.ctypes.data, len(bands), rows, rows)
            call("gm_bm_copy_to_device", C.byref(self.m), C.byref(h))
        self.bands = bands.copy()
        self.rows = rows

    @property
    def ref(self):
        return C.byref(self.m)

    def to_numpy(self, apitch: int | None = None) -> np.ndarray:
        ap = apitch or self.rows
        a = np.zeros((self.m.nbands, ap), np.float32)
        h = BandedMatrix(a.ctypes.data, ap, self.bands.ctypes.data, self.m.nbands, self.rows, self.rows)
        call("gm_bm_copy_to_host", C.byref(h), self.ref)
        return a

    def free(self):
        if self.m.a:
            call("gm_bm_device_free", self.ref)
            self.m.a = None

    def __del__(self):
        try:
            if self.m.a:
                lib().gm_bm_device_free(self.ref)
        except Exception:
            pass


def _dev(x, dtype=torch.float32):
    if x is None:
        return None
    t = torch.as_tensor(np.ascontiguousarray(x)) if isinstance(x, np.ndarray) else x
    return t.to(device="cuda", dtype=dtype).contiguous()


@dataclasses.dataclass
class System:
    L: DeviceMatrix
    b: torch.Tensor
    dinv: torch.Tensor
    W: int
    H: int

    @property
    def N(self):
        return self.W * self.H

    @property
    def pad(self):
        return levin_pad(self.W)


def assemble(image, lam=None, tgt=None, eps: float = 1e-7, mode: int = _lib.ASSEMBLY_ACCURATE,
             L: DeviceMatrix | None = None) -> System:
    """GPU Levin Laplacian (L + diag(lam)) and b = lam * tgt.  image: (H, W, 3|4)."""
    im = _dev(image)
    H, W, nch = im.shape
    N = W * H
    if L is None:
        L = DeviceMatrix(N, levin_bands(W))
    b = torch.empty(N, dtype=torch.float32, device="cuda")
    dinv = torch.empty(N, dtype=torch.float32, device="cuda")
    la = _dev(lam)
    tg = _dev(tgt)
    call("gm_levin_laplacian", L.ref, _p(b), _p(dinv), _p(im), nch, W, _p(la), _p(tg), eps, mode, W, H, _stream())
    return System(L, b, dinv, W, H)


def assemble_u8(image, scribs, eps: float = 1e-2, gamma: float = 1.0, mode: int = _lib.ASSEMBLY_REFERENCE) -> System:
    im = _dev(image)
    H, W, nch = im.shape
    L = DeviceMatrix(W * H, levin_bands(W))
    b = torch.empty(W * H, dtype=torch.float32, device="cuda")
    dinv = torch.empty(W * H, dtype=torch.float32, device="cuda")
    sc = _dev(scribs, torch.uint8)
    call("gm_levin_laplacian_u8", L.ref, _p(b), _p(dinv), _p(im), nch, W, _p(sc), eps, gamma, mode, W, H, _stream())
    return System(L, b, dinv, W, H)


def solve_opts(method=_lib.METHOD_PCG, precision=_lib.PREC_FP32, tol=1e-6, max_iter=10000, restart=0,
               refine_every=0, check_every=16, verify=True, refine_mode=0) -> SolveOpts:
    o = SolveOpts()
    lib().gm_solve_opts_default(C.byref(o))
    o.method, o.precision, o.tol, o.max_iter = method, precision, tol, max_iter
    o.restart, o.refine_every, o.check_every, o.verify = restart, refine_every, check_every, int(verify)
    o.refine_mode = refine_mode
    return o


def solve(sys_or_L, b=None, x: Padded | None = None, pad: int | None = None, raise_on_error=True, **kw):
    """gm_solve on a System (or a DeviceMatrix + b).  Returns (x, SolveInfo)."""
    if isinstance(sys_or_L, System):
        L, b, pad = sys_or_L.L, sys_or_L.b, sys_or_L.pad if pad is None else pad
    else:
        L = sys_or_L
    if pad is None:
        pad = int(np.abs(L.bands).max())
    if x is None:
        x = Padded(L.rows, pad)
    info = SolveInfo()
    o = solve_opts(**kw)
    st = lib().gm_solve(x.ptr, L.ref, _p(b), x.pad, C.byref(o), C.byref(info), _stream())
    if raise_on_error and st not in (0, 6):
        _lib.check(st)
    info.status = st
    return x, info


def true_relres(L: DeviceMatrix, x: Padded, b: torch.Tensor) -> float:
    r = C.c_double(0)
    call("gm_true_relres", L.ref, x.ptr, _p(b), C.byref(r), _stream())
    return r.value


def bm_ax(L: DeviceMatrix, x: Padded, y: torch.Tensor | None = None, add: bool | None = None) -> torch.Tensor:
    out = torch.empty(L.rows, dtype=torch.float32, device="cuda")
    if y is None:
        call("gm_bm_ax", _p(out), L.ref, x.ptr, _stream())
    else:
        call("gm_bm_axpy", _p(out), L.ref, x.ptr, _p(y), int(bool(add)), _stream())
    return out


def levin_backward(image, mu: torch.Tensor, alpha: torch.Tensor, lam=None, tgt=None, eps: float = 1e-7):
    """Assembly adjoint (gm_levin_backward): returns (d image (H,W,3), d lam (N), d tgt (N))."""
    im = _dev(image)
    H, W, nch = im.shape
    gim = torch.empty(H * W * 3, dtype=torch.float32, device="cuda")
    glam = torch.empty(H * W, dtype=torch.float32, device="cuda")
    gtg = torch.empty(H * W, dtype=torch.float32, device="cuda")
    la, tg = _dev(lam), _dev(tgt)
    call("gm_levin_backward", _p(gim), _p(glam), _p(gtg), _p(im), nch, W, _p(mu.contiguous()),
         _p(alpha.contiguous()), _p(la), _p(tg), eps, W, H, _stream())
    return gim.view(H, W, 3), glam, gtg


def autotune(L: DeviceMatrix, precision: int = _lib.PREC_FP32, reps: int = 5, full: bool = False):
    """gm_autotune: returns (rows_per_thread, tile_rows, wg_per_cu, step_us) and
    records it (full=True: the whole gm_tune_result)."""
    r = _lib.TuneResult()
    call("gm_autotune", L.ref, precision, reps, C.byref(r), _stream())
    if full:
        return {f: getattr(r, f) for f, _ in r._fields_}
    return r.rows_per_thread, r.tile_rows, r.wg_per_cu, r.step_us
