"""Driver I/O through the C ABI's host entry points (no GPU): Netpbm reading
and writing with ppm.h's normalisation and quantisation rules, computeError's
clamped MSE (main.cu:558-577) against the oracle, and the A.bin dump format
(main.cu:334-355)."""
import os

import numpy as np
import pytest

from gpumatting_amd import _lib, netpbm
from oracle import pyoracle as O
from tests.golden import netpbm as ref_reader

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "data05_crop.npz")


def _write_raw(path, magic, w, h, payload, maxval=255, header=None):
    with open(path, "wb") as f:
        f.write(header if header is not None else f"{magic}\n{w} {h}\n{maxval}\n".encode())
        f.write(payload)


def test_pgm_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    for binary in (True, False):
        p = tmp_path / f"x{int(binary)}.pgm"
        netpbm.write_pgm(p, a, comment="made here", binary=binary)
        got, m = netpbm.read_pgm(p)
        assert m == 255 and got.dtype == np.uint8 and np.array_equal(got, a)
    # header exactly as pgmwrite lays it out (ppm.h:367-384)
    raw = (tmp_path / "x1.pgm").read_bytes()
    assert raw.startswith(b"P5\n# made here \n53 37 \n255 \n") and len(raw) == len(b"P5\n# made here \n53 37 \n255 \n") + a.size
    netpbm.write_pgm(tmp_path / "y.pgm", a, comment=None)
    assert (tmp_path / "y.pgm").read_bytes().startswith(b"P5\n53 37 \n255 \n")


def test_ascii_and_binary_forms_agree(tmp_path):
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 200, (9, 11, 3), dtype=np.uint8)
    _write_raw(tmp_path / "b.ppm", "P6", 11, 9, rgb.tobytes(), maxval=200)
    txt = " ".join(str(v) for v in rgb.ravel()).encode()
    _write_raw(tmp_path / "t.ppm", "P3", 11, 9, txt, maxval=200,
               header=b"P3\n# a comment line\n11 9\n# another\n200\n")
    for name in ("b.ppm", "t.ppm"):
        got, m = netpbm.read_ppm(tmp_path / name)
        assert m == 200 and np.array_equal(got, rgb)
    f4 = netpbm.read_ppm_float4(tmp_path / "b.ppm")
    assert f4.shape == (9, 11, 4) and np.all(f4[..., 3] == 1.0)
    assert np.array_equal(f4[..., :3], rgb.astype(np.float32) / np.float32(200))


@pytest.mark.parametrize("header,data", [
    (b"P5 4 1 255 ", [66, 10, 32, 13]), (b"P5 4 1 255\t", [66, 32, 10, 13]),
    (b"P5\n4 1\n255\n", [10, 32, 66, 13]), (b"P5\n4 1\n255 \n", [10, 32, 66, 13]),
    (b"P5\n4 1\n255\r\n", [32, 10, 66, 13]), (b"P5\n# c\n4 1 # c2\n255\t \n", [13, 10, 32, 9])])
def test_sample_start_after_maxval(tmp_path, header, data):
    """One whitespace byte after maxval (the Netpbm rule), or blanks and a
    newline (what the reference's pgmwrite emits and its reader skips).
    Samples may look like whitespace; a blank after maxval followed by a
    sample of value 10 (a newline) is read the reference's way."""
    data = bytes(data)
    (tmp_path / "s.pgm").write_bytes(header + data)
    got, _ = netpbm.read_pgm(tmp_path / "s.pgm")
    assert got.ravel().tolist() == list(data)


def test_fixture_images_read_like_an_independent_reader(tmp_path):
    """The reference's own data/05 images (a crop kept as a fixture), written
    as binary Netpbm, read back identically by the library and by the
    fixture generator's separate reader."""
    d = np.load(GOLDEN)
    _write_raw(tmp_path / "im.ppm", "P6", 96, 64, d["rgb"].tobytes(), maxval=int(d["maxval"]))
    _write_raw(tmp_path / "sc.pgm", "P5", 96, 64, d["scribs"].tobytes())
    for name, key in (("im.ppm", "rgb"), ("sc.pgm", "scribs")):
        got, m = netpbm.read_ppm(tmp_path / name) if name.endswith("ppm") else netpbm.read_pgm(tmp_path / name)
        ref, rm = ref_reader.read(tmp_path / name)
        assert m == rm and np.array_equal(got, ref) and np.array_equal(got, d[key])
    g = netpbm.read_pgm_float(tmp_path / "sc.pgm")
    assert np.array_equal(g, d["scribs"].astype(np.float32) / np.float32(255))


@pytest.mark.parametrize("content,what", [
    (b"P4\n2 2\n255\n\0\0\0\0", "magic"),
    (b"P5\n2 2\n65535\n" + bytes(8), "maxval"),
    (b"P5\n2 2\n255\n\0\0\0", "short data"),
    (b"P5\n2\n", "header"),
    (b"P2\n2 2\n255\n1 2 3", "short ascii"),
    (b"P2\n2 2\n10\n1 2 3 11", "sample above maxval"),
    (b"", "empty"),
])
def test_bad_netpbm_files_are_rejected(tmp_path, content, what):
    p = tmp_path / "bad.pgm"
    p.write_bytes(content)
    with pytest.raises(_lib.GMError) as e:
        netpbm.read_pgm(p)
    assert e.value.status == 1, what
    with pytest.raises(_lib.GMError):
        netpbm.read_pgm(tmp_path / "missing.pgm")


def test_quantize_matches_pgmwrite_float_rule(tmp_path):
    v = np.array([-1.0, -1e-9, 0.0, 1 / 255, 0.5, 0.999, 1.0, 1.0 + 1e-7, 7.0, np.nan], np.float32)
    q = netpbm.quantize_u8(v)
    # v < 0 -> 0, v > 1 -> 255, else truncation of the FP32 product 255 * v
    want = [0, 0, 0, int(np.float32(255) * np.float32(1 / 255)), 127, 254, 255, 255, 255, 0]
    assert q.tolist() == want
    m = np.linspace(-0.2, 1.2, 60, dtype=np.float32).reshape(6, 10)
    netpbm.write_pgm_float(tmp_path / "m.pgm", m)
    got, _ = netpbm.read_pgm(tmp_path / "m.pgm")
    assert np.array_equal(got, netpbm.quantize_u8(m))
    assert (tmp_path / "m.pgm").read_bytes().startswith(b"P5\n#  \n10 6 \n255 \n")


def test_clamped_mse_matches_oracle():
    rng = np.random.default_rng(5)
    gt = rng.random((40, 30)).astype(np.float32)
    a = (rng.random((40, 30)) * 1.6 - 0.3).astype(np.float32)  # values outside [0, 1] too
    assert netpbm.clamped_mse(a, gt) == O.mse_clamped(a, gt)
    # closed forms: alpha above 1 scores as 1, below 0 as 0
    assert netpbm.clamped_mse(np.full_like(gt, 3.0), gt) == pytest.approx(np.mean((1.0 - gt.astype(np.float64)) ** 2))
    assert netpbm.clamped_mse(np.full_like(gt, -3.0), gt) == pytest.approx(np.mean(gt.astype(np.float64) ** 2))
    assert netpbm.clamped_mse(gt, gt) == 0.0


def test_matrix_dump_layout_and_round_trip(tmp_path):
    W, H = 13, 7
    N = W * H
    bands = O.bands25(W)
    rng = np.random.default_rng(2)
    apitch = N + 29  # padded host pitch: padding is not written
    a = rng.standard_normal((25, apitch)).astype(np.float32)
    netpbm.dump_matrix(tmp_path / "A.bin", a, bands, rows=N)
    raw = np.fromfile(tmp_path / "A.bin", np.float32)
    assert raw.size == 25 * N and np.array_equal(raw.reshape(25, N), a[:, :N])
    back = netpbm.load_matrix(tmp_path / "A.bin", bands, N, apitch=apitch)
    assert np.array_equal(back[:, :N], a[:, :N]) and not back[:, N:].any()
    with pytest.raises(_lib.GMError):
        netpbm.load_matrix(tmp_path / "A.bin", bands, N + 1)  # size mismatch
