"""CPU tests of the C-ABI library: it loads, exports every entry point that
include/gpumatting.h declares, and its host-only helpers and argument checks
work without a GPU (no compute call is made here)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from gpumatting_amd import _lib
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpumatting.h")).read()
    return sorted(set(re.findall(r"^(?:gm_status|void|int|const char\*)\s+(gm_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_symbols()
    assert len(decl) >= 30
    for name in decl:
        assert hasattr(L, name), name
    assert set(decl) == set(_lib.exported_symbols())
    nm = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in decl:
        assert re.search(r"\sT\s" + name + r"$", nm, re.M), name


def test_host_helpers_match_reference_layout():
    L = _lib.lib()
    for W in (3, 7, 256, 1024, 3840):
        b = np.zeros(25, np.int32)
        L.gm_levin_bands(W, b.ctypes.data)
        assert np.array_equal(b, O.bands25(W))
        assert L.gm_levin_pad(W) == O.pad_for(W) == ((2 * W + 2) // 32) * 32 + 32
    assert b[12] == 0 and b[0] == -(2 * 3840 + 2) and b[24] == 2 * 3840 + 2


def test_version_and_error_paths_without_gpu():
    L = _lib.lib()
    assert b"gfx950" in L.gm_version()
    # argument checks fire before any device call
    assert L.gm_bm_device_alloc(None, 10, 25, None) == 1
    assert b"bad arguments" in L.gm_last_error()
    m = _lib.BandedMatrix()
    assert L.gm_bm_ax(None, C.byref(m), None, None) == 1  # null matrix
    assert L.gm_solve(None, C.byref(m), None, 0, None, None, None) == 1
    assert L.gm_vec_add(None, None, None, -1, None) == 1
    bands = np.array([-1, 0, 1], np.int32)
    a = np.zeros((3, 4), np.float32)
    h = _lib.BandedMatrix(a.ctypes.data, 4, bands.ctypes.data, 3, 4, 4)
    x = np.zeros(6, np.float32)
    # jacobi_host with too little padding is rejected on the host
    assert L.gm_jacobi_host(x.ctypes.data + 4, C.byref(h), x.ctypes.data, 2, 0, C.c_float(0.5)) == 1
    # batched images of 2^25 pixels or more are rejected before any device call
    out = C.c_void_p()
    for w, h_ in ((16384, 16384), (1 << 20, 1 << 10), (8192, 4097), (0, 5)):
        W = np.array([64, w], np.int32)
        H = np.array([64, h_], np.int32)
        assert L.gm_batch_create(C.byref(out), 2, W.ctypes.data, H.ctypes.data, 4, 0) == 1
        assert b"gm_batch_create: image 1" in L.gm_last_error()
        assert not out.value
    W = np.array([16383], np.int32)
    H = np.array([16383], np.int32)
    assert L.gm_batch_create(C.byref(out), 1, W.ctypes.data, H.ctypes.data, 9, 0) == 1  # nrhs above 8


def test_cpp_dropin_headers_compile():
    """The reference-named C++ wrappers compile as plain C++ against the C ABI."""
    src = os.path.join(ROOT, "tests", "cpp", "dropin_compile.cpp")
    r = subprocess.run(["g++", "-std=c++14", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"),
                        "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_abi_header_is_plain_c():
    """include/gpumatting.h and the container headers compile as C99 with
    every warning an error, so C, cgo-style and ctypes callers can include them."""
    inc = os.path.join(ROOT, "include")
    for h in ("gpumatting.h", "BandedMatrix.h", "CompressedMatrix.h", "CoordMatrix.h"):
        r = subprocess.run(["gcc", "-x", "c", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", "-fsyntax-only",
                            "-I" + inc, os.path.join(inc, h)], capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr)


def test_ctypes_structs_match_the_header():
    """The ctypes mirrors of the option / info / tune structs have the C
    layout the header declares (field names, order, sizes)."""
    text = open(os.path.join(ROOT, "include", "gpumatting.h")).read()

    def fields(struct):
        body = re.search(r"typedef struct " + struct + r"\s*\{(.*?)\}", text, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        return [m.group(2) for m in re.finditer(r"\b(int|double|float)\s+(\w+)\s*;", body)]

    for struct, cls in (("gm_solve_opts_s", _lib.SolveOpts), ("gm_solve_info_s", _lib.SolveInfo),
                        ("gm_tune_result_s", _lib.TuneResult)):
        assert fields(struct) == [f for f, _ in cls._fields_], struct


def test_sparse_structs_match_the_headers():
    """The ctypes mirrors of CompressedMatrix / CoordMatrix have the headers'
    field names and order (the reference's, CompressedMatrix.h:26-37,
    CoordMatrix.h:25-36)."""
    for h, cls in (("CompressedMatrix.h", _lib.CompressedMatrix), ("CoordMatrix.h", _lib.CoordMatrix)):
        text = open(os.path.join(ROOT, "include", h)).read()
        body = re.search(r"typedef struct \w+\s*\{(.*?)\}", text, re.S).group(1)
        names = [m.group(1) for m in re.finditer(r"(\w+)\s*;", body)]
        assert names == [f for f, _ in cls._fields_], h
