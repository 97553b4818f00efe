"""The solver's float -> FP16 conversion (gpumatting_amd/csrc/gm_half.h, the
integer round-to-nearest-even the 16-bit operator is stored with), built for
the host, against numpy's float32 -> float16 conversion on every float32 bit
pattern: ties to even, subnormal halves, overflow to infinity, signed zeros.
NaN inputs must give a NaN (numpy keeps payload bits, the solver a quiet NaN)."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "cpp", "bin", "libhalfconvert.so")


@pytest.fixture(scope="module")
def conv():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (make -C tests/cpp)")
    lib = C.CDLL(LIB)
    lib.half_convert_range.argtypes = [C.c_uint32, C.c_size_t, C.c_void_p]
    return lib.half_convert_range


def test_known_values(conv):
    out = np.empty(1, np.uint16)
    for bits, want in [(0x3F800000, 0x3C00),   # 1
                       (0x3DA8D000, 0x2D46),   # a float that is an FP16 tie: to even
                       (0x3DA8F000, 0x2D48),   # the next tie rounds up to even
                       (0x477FE000, 0x7BFF),   # 65504
                       (0x477FF000, 0x7C00),   # 65520 -> inf
                       (0x33800000, 0x0001),   # 2^-24, smallest subnormal
                       (0x33000000, 0x0000),   # 2^-25 ties to 0
                       (0x33000001, 0x0001),
                       (0x387FC000, 0x03FF),   # largest subnormal
                       (0x387FE000, 0x0400),   # rounds up into the normals
                       (0x80000000, 0x8000)]:  # -0
        conv(bits, 1, out.ctypes.data)
        assert out[0] == want, hex(bits)


def test_every_nonnegative_float32(conv):
    """All 2^31 patterns with the sign bit clear up to +inf, exhaustively:
    numpy's conversion where the result is a rounding (2^-25 .. 65520), the
    constants 0 and +inf below and above (numpy's per-element underflow and
    overflow flags make it slow there)."""
    chunk = 1 << 24
    out = np.empty(chunk, np.uint16)
    lo, hi = 0x33000000, 0x477FF000  # 2^-25 and 65520
    for first in range(0, 0x7F800000 + 1, chunk):
        n = min(chunk, 0x7F800001 - first)  # stop after +inf
        conv(first, n, out.ctypes.data)
        h = out[:n]
        if first + n <= lo:
            assert not h.any(), hex(first)
        elif first >= hi:
            assert np.all(h[: min(n, 0x7F800000 - first)] == 0x7C00) and h[-1] == 0x7C00, hex(first)
        else:
            bits = (np.arange(n, dtype=np.uint64) + first).astype(np.uint32)
            with np.errstate(over="ignore", under="ignore"):
                ref = bits.view(np.float32).astype(np.float16).view(np.uint16)
            bad = np.nonzero(h != ref)[0]
            assert bad.size == 0, f"{int(bits[bad[0]]):#010x}: {int(h[bad[0]]):#06x} vs {int(ref[bad[0]]):#06x}"


def test_sign_and_nan_samples(conv):
    """Negative patterns (windows strided through all of them) mirror the
    positive ones with the sign bit set; every NaN pattern sampled gives a NaN."""
    n = 1 << 20
    out = np.empty(n, np.uint16)
    pos = np.empty(n, np.uint16)
    for first in range(0x80000000, (1 << 32) - n, (1 << 31) // 37):
        conv(first, n, out.ctypes.data)
        conv(first - 0x80000000, n, pos.ctypes.data)
        f = (np.arange(n, dtype=np.uint64) + first).astype(np.uint32).view(np.float32)
        nan = np.isnan(f)
        assert np.all(((out[nan] & 0x7C00) == 0x7C00) & ((out[nan] & 0x03FF) != 0))
        assert np.array_equal(out[~nan], pos[~nan] | 0x8000), hex(first)
    for first in (0x7F800001, 0x7FC00000, 0x7FFFFFFF - n):
        conv(first, n, out.ctypes.data)
        assert np.all(((out & 0x7C00) == 0x7C00) & ((out & 0x03FF) != 0)), hex(first)
