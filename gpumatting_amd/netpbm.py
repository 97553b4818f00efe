"""Image and matrix file I/O of the matting driver, over the library's host
entry points (gm_pgm_read, gm_ppm_read, gm_pgm_write[_float],
gm_quantize_u8, gm_clamped_mse, gm_bm_dump, gm_bm_load): the reference's
ppm.h readers / writers, main.cu's computeError and its A.bin dump
(ppm.h:32-451, main.cu:334-355, 558-577).  No device work."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import BandedMatrix, call, lib


def _path(p) -> bytes:
    return os.fsencode(p)


def _take(ptr: C.c_void_p, n: int) -> np.ndarray:
    try:
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), (n,)).copy()
    finally:
        lib().gm_free(ptr)


def read_pgm(path) -> tuple[np.ndarray, int]:
    """(H, W) uint8 samples and maxval of a P5 / P2 file (ppm.h:32-129)."""
    d, w, h, m = C.c_void_p(), C.c_int(), C.c_int(), C.c_int()
    call("gm_pgm_read", _path(path), C.byref(d), C.byref(w), C.byref(h), C.byref(m))
    return _take(d, w.value * h.value).reshape(h.value, w.value), m.value


def read_ppm(path) -> tuple[np.ndarray, int]:
    """(H, W, 3) uint8 samples and maxval of a P6 / P3 file (ppm.h:178-278)."""
    d, w, h, m = C.c_void_p(), C.c_int(), C.c_int(), C.c_int()
    call("gm_ppm_read", _path(path), C.byref(d), C.byref(w), C.byref(h), C.byref(m))
    return _take(d, 3 * w.value * h.value).reshape(h.value, w.value, 3), m.value


def read_pgm_float(path) -> np.ndarray:
    """pgmread_float (ppm.h:134-149): samples / 255."""
    a, _ = read_pgm(path)
    return a.astype(np.float32) / np.float32(255.0)


def read_ppm_float4(path) -> np.ndarray:
    """ppmread_float4 (ppm.h:309-337): (H, W, 4) RGBA, samples / maxval, alpha 1."""
    a, m = read_ppm(path)
    out = np.ones(a.shape[:2] + (4,), np.float32)
    out[..., :3] = a.astype(np.float32) / np.float32(m)
    return out


def write_pgm(path, data: np.ndarray, comment: str | None = None, binary: bool = True):
    """pgmwrite (ppm.h:345-418) of (H, W) uint8 samples."""
    d = np.ascontiguousarray(data, np.uint8)
    h, w = d.shape
    call("gm_pgm_write", _path(path), d.ctypes.data, w, h, None if comment is None else comment.encode(), int(binary))


def quantize_u8(v: np.ndarray) -> np.ndarray:
    """pgmwrite_float's quantisation (ppm.h:430-451)."""
    f = np.ascontiguousarray(v, np.float32)
    out = np.empty(f.shape, np.uint8)
    lib().gm_quantize_u8(f.ctypes.data, f.size, out.ctypes.data)
    return out


def write_pgm_float(path, data: np.ndarray, comment: str | None = "", binary: bool = True):
    """pgmwrite_float (ppm.h:430-451) of an (H, W) float matte."""
    f = np.ascontiguousarray(data, np.float32)
    h, w = f.shape
    call("gm_pgm_write_float", _path(path), f.ctypes.data, w, h, None if comment is None else comment.encode(),
         int(binary))


def clamped_mse(alpha: np.ndarray, gt: np.ndarray) -> float:
    """computeError (main.cu:558-577)."""
    a = np.ascontiguousarray(alpha, np.float32).ravel()
    g = np.ascontiguousarray(gt, np.float32).ravel()
    if a.size != g.size:
        raise ValueError("alpha and gt differ in size")
    out = C.c_double()
    call("gm_clamped_mse", a.ctypes.data, g.ctypes.data, a.size, C.byref(out))
    return out.value


def _host_bm(a: np.ndarray, bands: np.ndarray, rows: int) -> BandedMatrix:
    return BandedMatrix(a.ctypes.data, a.shape[1], bands.ctypes.data, a.shape[0], rows, rows)


def dump_matrix(path, a: np.ndarray, bands, rows: int | None = None):
    """dump (main.cu:334-355): bands (nbands, apitch) float32, rows <= apitch."""
    a = np.ascontiguousarray(a, np.float32)
    bands = np.ascontiguousarray(bands, np.int32)
    m = _host_bm(a, bands, a.shape[1] if rows is None else rows)
    call("gm_bm_dump", C.byref(m), _path(path))


def load_matrix(path, bands, rows: int, apitch: int | None = None) -> np.ndarray:
    """Read an A.bin dump of a matrix with these bands and rows: (nbands, apitch)."""
    bands = np.ascontiguousarray(bands, np.int32)
    a = np.zeros((len(bands), apitch or rows), np.float32)
    m = _host_bm(a, bands, rows)
    call("gm_bm_load", C.byref(m), _path(path))
    return a
