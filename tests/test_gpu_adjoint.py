"""GPU differentiation path: the assembly adjoint kernel vs the FP64 oracle
(same mu, alpha), and the autograd wrapper's gradients vs the FP64 host
reference gradient (forward solve + adjoint solve + assembly adjoint)."""
import numpy as np
import pytest
import torch

from gpumatting_amd import autograd, matting, synth
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,H", [(7, 6), (33, 9), (64, 48), (512, 512)])
def test_assembly_adjoint_vs_oracle(gpu, W, H):
    p = synth.make_problem(W, H, 40 + W, 1, 0.08)
    rng = np.random.default_rng(W)
    mu = rng.standard_normal(p.N).astype(np.float32)
    al = rng.random(p.N).astype(np.float32)
    gim, glam, gtg = matting.levin_backward(p.image, torch.from_numpy(mu).cuda(), torch.from_numpy(al).cuda(),
                                            p.lam, p.target, 1e-7)
    ref = O.levin_backward_f64(p.image, mu.astype(np.float64), al.astype(np.float64), 1e-7)
    got = gim.cpu().numpy().astype(np.float64)
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() <= 1e-5 * scale
    assert np.allclose(glam.cpu().numpy(), mu * (p.target.ravel() - al), rtol=1e-6, atol=1e-6)
    assert np.allclose(gtg.cpu().numpy(), mu * p.lam.ravel(), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("W,H,seed", [(48, 40, 1), (96, 64, 2), (512, 512, 5000)])
def test_autograd_vs_fp64_reference_gradient(gpu, W, H, seed):
    """Gradients of loss = ||alpha - alpha_gt||^2 w.r.t. the RGB pixels, the
    per-pixel weights and the targets vs the FP64 host reference gradient
    (whose adjoint the oracle tests pin against central differences); the
    512x512 case is the first image of the differentiation workload."""
    p = synth.make_problem(W, H, seed, 1, 0.08)
    im = torch.from_numpy(p.image).cuda().requires_grad_(True)
    lam = torch.from_numpy(p.lam).cuda().requires_grad_(True)
    tgt = torch.from_numpy(p.target).cuda().requires_grad_(True)
    gt = torch.from_numpy(p.alpha_gt).cuda()
    alpha = autograd.matte(im, lam, tgt, 1e-7, 1e-9)
    loss = ((alpha - gt) ** 2).sum()
    loss.backward()
    L64, a64, gim, glam, gtg = O.matte_loss_grad_f64(p.image, p.lam, p.target, p.alpha_gt, 1e-7)
    assert abs(loss.item() - L64) <= 1e-4 * L64
    for got, ref in ((im.grad, gim), (lam.grad, glam.reshape(H, W)), (tgt.grad, gtg.reshape(H, W))):
        g = got.detach().cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        assert rel <= 1e-4, rel


def test_functional_gradient_equals_autograd(gpu):
    """matte_loss_and_grad makes the autograd wrapper's calls without a graph:
    the same loss and gradients, also from a side stream."""
    p = synth.make_problem(96, 64, 9, 1, 0.08)
    im = torch.from_numpy(p.image).cuda().requires_grad_(True)
    lam = torch.from_numpy(p.lam).cuda().requires_grad_(True)
    tgt = torch.from_numpy(p.target).cuda().requires_grad_(True)
    gt = torch.from_numpy(p.alpha_gt).cuda()
    alpha = autograd.matte(im, lam, tgt, 1e-7, 1e-9)
    loss = ((alpha - gt) ** 2).sum()
    loss.backward()
    st = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(st):
        out = autograd.matte_loss_and_grad(im.detach(), lam.detach(), tgt.detach(), gt, 1e-7, 1e-9)
    st.synchronize()
    l2, a2, g_im, g_lam, g_tgt = out
    assert torch.equal(a2, alpha.detach())
    assert abs(l2.item() - loss.item()) <= 1e-6 * loss.item()
    for got, ref in ((g_im, im.grad), (g_lam, lam.grad), (g_tgt, tgt.grad)):
        assert torch.equal(got, ref)
