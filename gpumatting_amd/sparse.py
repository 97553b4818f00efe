"""Sparse matrices (CSR, COO) and the superpixel multilevel operators, over
the C ABI (include/gpumatting.h, "Sparse matrices and superpixel multilevel
operators").  Matrices own device arrays allocated by the library; vectors
are CUDA float32 tensors.  Every computation is a library call.

Reference interfaces: CompressedMatrix.h/.cu (CSR), CoordMatrix.h/.cu (COO),
gridoperators.h (downsampleOps), SLIC.h:43 (slicSegmentation).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import CompressedMatrix as _CSM, CoordMatrix as _COO, call, lib
from .matting import DeviceMatrix, _p, _stream


def _vec(x, n: int, what: str) -> torch.Tensor:
    """x as a contiguous CUDA float32 vector of n elements (the library reads
    exactly n; a shorter buffer would be read past its end)."""
    t = torch.as_tensor(np.ascontiguousarray(x, np.float32)) if isinstance(x, np.ndarray) else x
    t = t.to(device="cuda", dtype=torch.float32).contiguous().reshape(-1)
    if t.numel() != n:
        raise ValueError(f"{what}: {t.numel()} elements, expected {n}")
    return t


def _host_copy(ptr, n: int, dtype) -> np.ndarray:
    out = np.empty(n, dtype)
    if n:
        call("gm_vec_copy_to_host", out.ctypes.data, C.c_void_p(ptr), n)
    return out


class CompressedMatrix:
    """A device CSR matrix (csmInit, CompressedMatrix.cu:20-31)."""

    def __init__(self, rows: int = 0, cols: int = 0, length: int = 0, _raw: _CSM | None = None):
        self.m = _raw if _raw is not None else _CSM()
        if _raw is None:
            call("gm_csm_init", C.byref(self.m), rows, cols, length)

    rows = property(lambda self: self.m.rows)
    cols = property(lambda self: self.m.cols)
    nnz = property(lambda self: self.m.nnz)

    @property
    def ref(self):
        return C.byref(self.m)

    def ax(self, x: torch.Tensor) -> torch.Tensor:
        """csmAx: A x."""
        x = _vec(x, self.cols, "x")
        b = torch.empty(self.rows, dtype=torch.float32, device="cuda")
        call("gm_csm_ax", _p(b), self.ref, _p(x), _stream())
        return b

    def axpy(self, x: torch.Tensor, y: torch.Tensor, add: bool = True) -> torch.Tensor:
        """csmAxpy: A x + y (add) or A x - y."""
        x, y = _vec(x, self.cols, "x"), _vec(y, self.rows, "y")
        b = torch.empty(self.rows, dtype=torch.float32, device="cuda")
        call("gm_csm_axpy", _p(b), self.ref, _p(x), _p(y), int(add), _stream())
        return b

    def to_numpy(self):
        """(p, j, k) as host arrays."""
        torch.cuda.synchronize()
        p = _host_copy(self.m.p, self.rows + 1, np.int32)
        j = _host_copy(self.m.j, self.nnz, np.int32)
        k = _host_copy(self.m.k, self.nnz, np.float32)
        return p, j, k

    def to_scipy(self):
        import scipy.sparse as sps
        p, j, k = self.to_numpy()
        return sps.csr_matrix((k, j, p), shape=(self.rows, self.cols))

    def free(self):
        if self.m.p:
            call("gm_csm_free", self.ref)

    def __del__(self):
        try:
            if self.m.p:
                lib().gm_csm_free(C.byref(self.m))
        except Exception:
            pass


class CoordMatrix:
    """A device COO matrix (cooInit / cooAppend / cooRelease, CoordMatrix.cu:20-48)."""

    def __init__(self, rows: int, cols: int, length: int):
        self.m = _COO()
        call("gm_coo_init", C.byref(self.m), rows, cols, length)

    rows = property(lambda self: self.m.rows)
    cols = property(lambda self: self.m.cols)
    nnz = property(lambda self: self.m.nnz)

    @property
    def ref(self):
        return C.byref(self.m)

    def append(self, i, j, k):
        i = np.ascontiguousarray(i, np.int32)
        j = np.ascontiguousarray(j, np.int32)
        k = np.ascontiguousarray(k, np.float32)
        if not (len(i) == len(j) == len(k)):
            raise ValueError("i, j, k differ in length")
        call("gm_coo_append", self.ref, i.ctypes.data, j.ctypes.data, k.ctypes.data, len(i))

    def to_csr(self) -> CompressedMatrix:
        out = CompressedMatrix(_raw=_CSM())
        call("gm_coo_to_csm", out.ref, self.ref, _stream())
        return out

    def axpy_(self, b: torch.Tensor, x: torch.Tensor, y: torch.Tensor) -> torch.Tensor:
        """cooAxpy: b += A x + y, in place (b: a contiguous CUDA float32 vector)."""
        if not (b.is_cuda and b.dtype == torch.float32 and b.is_contiguous() and b.numel() == self.rows):
            raise ValueError("b must be a contiguous CUDA float32 vector of `rows` elements")
        x, y = _vec(x, self.cols, "x"), _vec(y, self.rows, "y")
        call("gm_coo_axpy", _p(b), self.ref, _p(x), _p(y), _stream())
        return b

    def release(self):
        if self.m.i or self.m.length:
            call("gm_coo_release", self.ref)

    def __del__(self):
        try:
            if self.m.i:
                lib().gm_coo_release(C.byref(self.m))
        except Exception:
            pass


def _labels(labels) -> torch.Tensor:
    t = torch.as_tensor(np.ascontiguousarray(labels, np.int32)) if isinstance(labels, np.ndarray) else labels
    return t.to(device="cuda", dtype=torch.int32).contiguous().reshape(-1)


def restriction(labels, nlabels: int) -> CompressedMatrix:
    """P (nlabels x npix) with P[labels[u], u] = 1 (downsampleOps, first level)."""
    lt = _labels(labels)
    out = CompressedMatrix(_raw=_CSM())
    call("gm_sp_restriction", out.ref, _p(lt), lt.numel(), nlabels, _stream())
    return out


def prolong(xc: torch.Tensor, labels, nlabels: int) -> torch.Tensor:
    """P^T xc: each pixel takes its superpixel's value."""
    lt = _labels(labels)
    xc = _vec(xc, nlabels, "xc")
    e = torch.empty(lt.numel(), dtype=torch.float32, device="cuda")
    call("gm_sp_prolong", _p(e), _p(xc), _p(lt), lt.numel(), nlabels, _stream())
    return e


def galerkin(A: DeviceMatrix, labels, nlabels: int) -> CompressedMatrix:
    """The coarse operator P A P^T of a device banded matrix."""
    lt = _labels(labels)
    if lt.numel() != A.rows:
        raise ValueError(f"labels: {lt.numel()} elements, the matrix has {A.rows} rows")
    out = CompressedMatrix(_raw=_CSM())
    call("gm_sp_galerkin", out.ref, A.ref, _p(lt), nlabels, _stream())
    return out


def merge_levels(labels, nlabels: int, image, nlevels: int, threshold: float = 0.0111):
    """downsampleOps levels 2..nlevels (host): returns (labels, nlabels)."""
    im = np.ascontiguousarray(image, np.float32)
    H, W, nch = im.shape
    lab = np.ascontiguousarray(labels, np.int32).reshape(-1).copy()
    if lab.size != H * W:
        raise ValueError(f"labels: {lab.size} elements for a {W}x{H} image")
    n = C.c_int(nlabels)
    call("gm_sp_merge_levels", lab.ctypes.data, C.byref(n), im.ctypes.data, nch, W, W, H, nlevels, threshold)
    return lab.reshape(H, W), n.value


def slic(image, K: int, compactness: float = 10.0, iters: int = 10):
    """SLIC superpixels of an (H, W, 3|4) RGB image in [0, 1]: (labels (H, W)
    int32 CUDA tensor, number of superpixels)."""
    t = torch.as_tensor(np.ascontiguousarray(image, np.float32)) if isinstance(image, np.ndarray) else image
    im = t.to(device="cuda", dtype=torch.float32).contiguous()
    H, W, nch = im.shape
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    n = C.c_int(0)
    call("gm_slic", _p(labels), C.byref(n), _p(im), nch, W, W, H, K, compactness, iters, _stream())
    return labels, n.value


def slic_grid(W: int, H: int, K: int):
    """(S, nsx, nsy): the grid step and seed counts SLIC uses for K superpixels."""
    S, nx, ny = C.c_int(), C.c_int(), C.c_int()
    call("gm_slic_grid", W, H, K, C.byref(S), C.byref(nx), C.byref(ny))
    return S.value, nx.value, ny.value


def slic_lab(image) -> torch.Tensor:
    """The CIELAB planes SLIC clusters: (3, H, W) float32 CUDA tensor."""
    t = torch.as_tensor(np.ascontiguousarray(image, np.float32)) if isinstance(image, np.ndarray) else image
    im = t.to(device="cuda", dtype=torch.float32).contiguous()
    H, W, nch = im.shape
    lab = torch.empty((3, H, W), dtype=torch.float32, device="cuda")
    call("gm_slic_lab", _p(lab), _p(im), nch, W, W, H, _stream())
    return lab


def slic_raw(lab, K: int, compactness: float = 10.0, iters: int = 10):
    """Seeds and `iters` rounds from Lab planes (3, H, W): (raw labels (H, W)
    int32, seeds (5, nseeds) float32: l, a, b, x, y), both CUDA tensors."""
    t = torch.as_tensor(np.ascontiguousarray(lab, np.float32)) if isinstance(lab, np.ndarray) else lab
    lab = t.to(device="cuda", dtype=torch.float32).contiguous()
    _, H, W = lab.shape
    _, nx, ny = slic_grid(W, H, K)
    labels = torch.empty((H, W), dtype=torch.int32, device="cuda")
    seeds = torch.empty((5, nx * ny), dtype=torch.float32, device="cuda")
    call("gm_slic_raw", _p(labels), _p(seeds), nx * ny, _p(lab), W, H, K, compactness, iters, _stream())
    return labels, seeds


def slic_connect(labels, minsize: int):
    """Connectivity of raw labels (H, W): (labels (H, W) int32 CUDA tensor,
    number of superpixels)."""
    t = torch.as_tensor(np.ascontiguousarray(labels, np.int32)) if isinstance(labels, np.ndarray) else labels
    out = t.to(device="cuda", dtype=torch.int32).contiguous().clone()
    H, W = out.shape
    n = C.c_int(0)
    call("gm_slic_connect", _p(out), C.byref(n), W, H, minsize, _stream())
    return out, n.value
