"""The batched solver's memory accesses, modelled on the CPU
(tests/cpp/batch_model): every slot, LDS ring, vector and fold index its
kernels form, with the kernels' own index functions, over edge-case image
sizes and images of the 64-image batched workload, at several grids and every
tile width; and that each pixel is advanced and output exactly once per
launch."""
import os
import shutil
import subprocess

import pytest

from gpumatting_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "batch_model", "batch_index_model.cpp")


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("bim") / "batch_index_model"
    b = subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), SRC], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    return str(exe)


def _run(model, nrhs, per_cu, sizes):
    args = [model, str(nrhs), "256", *map(str, per_cu), "--"] + [str(v) for wh in sizes for v in wh]
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout, r.stdout


EDGE = [(1, 1), (1, 300), (300, 1), (2, 2), (63, 5), (64, 5), (65, 5), (257, 3), (128, 2), (129, 7), (5, 40),
        (17, 9), (256, 1), (1000, 3), (3, 1000)]


@pytest.mark.parametrize("nrhs", [1, 3, 8])
def test_edge_sizes(model, nrhs):
    _run(model, nrhs, [1, 2, 16, 64], EDGE)


def test_batched_workload_images(model):
    """The largest, smallest, widest and tallest images of the workload and two
    more, in one batch, at 1, 16 and 32 workgroups per CU (one range spanning
    several images, a tuner candidate, and the default grid)."""
    plan = synth.config_batched_plan()
    n = [w * h for w, h, _ in plan]
    asp = [w / h for w, h, _ in plan]
    pick = sorted({n.index(max(n)), n.index(min(n)), asp.index(max(asp)), asp.index(min(asp)), 0, 33})
    _run(model, 4, [1, 16, 32], [plan[i][:2] for i in pick])


def test_range_map_is_a_bijection(tmp_path):
    """range_of (workgroup -> range of a batched launch) is a bijection for
    every grid size up to 20000, gives each XCD an equal share, and keeps each
    block of consecutive ranges on one XCD."""
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = os.path.join(HERE, "cpp", "batch_model", "range_map_check.cpp")
    exe = tmp_path / "range_map_check"
    b = subprocess.run(["g++", "-O2", "-std=c++17", "-o", str(exe), src], capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr
    r = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failed" in r.stdout, r.stdout

