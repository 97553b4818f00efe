"""Batched solver on the batched workload at its own sizes: 64 images with
W ~ U[240,1080] and H ~ U[320,1920], 4 constraint sets (densities U[2 %,10 %])
per image, every system solved to 1e-6 relative residual.

Two tests:
* five images at full size -- the largest, the smallest, the widest and the
  tallest of the 64, plus one at the distribution's corner (1080x1920) --
  with all 4 right-hand sides in FP32, FP16 and BF16.  Every system is checked
  by the solver's own FP64 true residual, by a host FP64 residual of the
  returned alpha, and against the oracle's FP64 Jacobi-PCG solution.
* the whole 64 x 4 workload in every precision, each of the 256 systems'
  residual recomputed on the host in FP64.

The host residual is oracle.resid_data_f64 (pinned to scipy.sparse in
test_oracle.py); the Laplacians are the oracle's accurate-mode assembly, which
the GPU's batched assembly equals bit for bit (test_gpu_core.py).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from gpumatting_amd import PREC_BF16, PREC_FP16, PREC_FP32, synth
from gpumatting_amd.batch import Batch
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

PRECS = [("fp32", PREC_FP32), ("fp16", PREC_FP16), ("bf16", PREC_BF16)]
TOL = 1e-6


def _picks(plan):
    """The largest, smallest, widest (W/H) and tallest (H/W) images of the plan."""
    n = [w * h for w, h, _ in plan]
    asp = [w / h for w, h, _ in plan]
    out = []
    for i in (int(np.argmax(n)), int(np.argmin(n)), int(np.argmax(asp)), int(np.argmin(asp))):
        if i not in out:
            out.append(i)
    return out


def _generate(items):
    """Build (problem, rhs) for each (index, W, H, densities) on a thread pool."""
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda it: synth.batched_image(*it), items))


def _solve(work, prec, tol=TOL):
    b = Batch([(p.W, p.H) for p, _ in work], len(work[0][1]), prec)
    for i, (p, rhs) in enumerate(work):
        b.assemble(i, p.image, 1e-7)
        for j, (lam, tgt) in enumerate(rhs):
            b.set_rhs(i, j, lam, tgt)
    b.solve(tol=tol, max_iter=20000)
    return b


@pytest.fixture(scope="module")
def five_images():
    plan = synth.config_batched_plan()
    picks = _picks(plan)
    assert plan[picks[0]][0] * plan[picks[0]][1] >= 1_500_000, "the largest image of the workload"
    items = [(i, *plan[i]) for i in picks] + [(64, 1080, 1920, [0.02, 0.045, 0.075, 0.10])]
    work = _generate(items)
    ref = []  # per image: (Laplacian without data term, [FP64 solution per rhs])
    for p, rhs in work:
        a0, _, _ = O.levin_gather_acc_f32(p.image, None, None, 1e-7)
        xs = []
        for lam, tgt in rhs:
            a, bb, _ = O.levin_gather_acc_f32(p.image, lam, tgt, 1e-7)
            xe, _, _ = O.pcg_f64(a.astype(np.float64), O.bands25(p.W), bb.astype(np.float64), np.zeros(p.N),
                                 O.pad_for(p.W), 1e-10)
            xs.append(xe)
            del a
        ref.append((a0, xs))
    return work, ref


def test_batched_config_sizes_all_precisions(gpu, five_images, capsys):
    work, ref = five_images
    for name, prec in PRECS:
        b = _solve(work, prec)
        for i, (p, rhs) in enumerate(work):
            a0, xs = ref[i]
            for j, (lam, tgt) in enumerate(rhs):
                inf = b.info(i, j)
                assert inf.status == 0 and inf.true_relres <= TOL, (name, i, j, inf.status, inf.true_relres)
                x = b.alpha(i, j).cpu().numpy()
                res, bn = O.resid_data_f64(a0, O.bands25(p.W), lam, tgt, x, O.pad_for(p.W))
                assert res <= TOL * bn * 1.001, (name, i, j, res / bn)
                # the solver's FP64 residual and the host's describe the same vector
                assert abs(res / bn - inf.true_relres) <= 1e-3 * inf.true_relres + 1e-12, (name, i, j)
                d = np.abs(x.astype(np.float64) - xs[j])
                with capsys.disabled():
                    print(f"\n[batched {name}] image {i} {p.W}x{p.H} rhs {j}: {inf.iterations} it, "
                          f"relres {res / bn:.2e}, max|da| {d.max():.2e}, mean|da| {d.mean():.2e}", end="")
                assert d.max() <= 5e-2 and d.mean() <= 1e-4, (name, i, j, d.max(), d.mean())
        b.close()


def test_batched_full_workload_host_residuals(gpu, capsys):
    plan = synth.config_batched_plan()
    work = _generate([(i, *plan[i]) for i in range(len(plan))])
    assert len(work) == 64 and all(len(r) == 4 for _, r in work)
    batches = [(name, _solve(work, prec)) for name, prec in PRECS]
    worst = {name: 0.0 for name, _ in PRECS}
    for i, (p, rhs) in enumerate(work):
        a0, _, _ = O.levin_gather_acc_f32(p.image, None, None, 1e-7)
        for name, b in batches:
            for j, (lam, tgt) in enumerate(rhs):
                inf = b.info(i, j)
                assert inf.status == 0 and inf.true_relres <= TOL, (name, i, j, inf.status, inf.true_relres)
                x = b.alpha(i, j).cpu().numpy()
                res, bn = O.resid_data_f64(a0, O.bands25(p.W), lam, tgt, x, O.pad_for(p.W))
                assert res <= TOL * bn * 1.001, (name, i, j, res / bn)
                worst[name] = max(worst[name], res / bn)
        del a0
    with capsys.disabled():
        print("\n[batched 64x4] worst host FP64 relative residual: "
              + ", ".join(f"{k} {v:.3e}" for k, v in worst.items()))
    for _, b in batches:
        b.close()
