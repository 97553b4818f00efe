"""torch.autograd wrapper of the closed-form matte (for gradient checks).

    alpha = LevinMatte.apply(image, lam, tgt, eps, tol)

Forward: GPU assembly (accurate rounding) of (L(image) + diag(lam)), then
Jacobi-PCG in FP32 with an FP64 master solution and FP64 residual
replacement (refine_mode 2; 1 selects outer iterative refinement) to `tol`.
Backward: one adjoint solve mu = (L + diag(lam))^-1 grad_alpha (same solver;
the system is symmetric) and the assembly adjoint kernel, which returns the
gradients with respect to the RGB pixels, the per-pixel weights lam and the
targets.  All arithmetic runs in libgpumatting; torch only holds the tensors.
"""
from __future__ import annotations

import torch

from . import _lib, matting


def _solve(s: matting.System, rhs: torch.Tensor, tol: float, max_iter: int, refine_mode: int):
    x = matting.Padded(s.N, s.pad)
    _, info = matting.solve(s.L, rhs.reshape(-1).contiguous(), x=x, method=_lib.METHOD_PCG,
                            precision=_lib.PREC_FP32, tol=tol, max_iter=max_iter, refine_mode=refine_mode)
    if info.status != 0:
        raise _lib.GMError(info.status, f"matte solve did not converge (relres {info.relres:.3e})")
    return x.t.clone(), info


class LevinMatte(torch.autograd.Function):
    @staticmethod
    def forward(ctx, image, lam, tgt, eps=1e-7, tol=1e-8, max_iter=20000, refine_mode=2):
        H, W = image.shape[:2]
        s = matting.assemble(image.detach(), lam.detach(), tgt.detach(), eps, _lib.ASSEMBLY_ACCURATE)
        alpha, info = _solve(s, s.b, tol, max_iter, refine_mode)
        ctx.system = s
        ctx.eps, ctx.tol, ctx.max_iter, ctx.refine_mode = eps, tol, max_iter, refine_mode
        ctx.save_for_backward(image, lam, tgt, alpha)
        return alpha.view(H, W)

    @staticmethod
    def backward(ctx, grad_alpha):
        image, lam, tgt, alpha = ctx.saved_tensors
        s = ctx.system
        mu, _ = _solve(s, grad_alpha.reshape(-1).float(), ctx.tol, ctx.max_iter, ctx.refine_mode)
        gim, glam, gtg = matting.levin_backward(image.detach(), mu, alpha, lam.detach(), tgt.detach(), ctx.eps)
        H, W = image.shape[:2]
        gi = torch.zeros_like(image)
        gi[..., :3] = gim.to(image.dtype)
        return gi, glam.view_as(lam).to(lam.dtype), gtg.view_as(tgt).to(tgt.dtype), None, None, None, None


def matte(image, lam, tgt, eps=1e-7, tol=1e-8, max_iter=20000, refine_mode=2):
    return LevinMatte.apply(image, lam, tgt, eps, tol, max_iter, refine_mode)


def matte_loss_and_grad(image, lam, tgt, alpha_gt, eps=1e-7, tol=1e-8, max_iter=20000, refine_mode=2):
    """loss = ||alpha - alpha_gt||^2 and its gradients with respect to the
    image, lam and tgt without an autograd graph: the same forward solve,
    adjoint solve and assembly adjoint LevinMatte runs, on the calling thread's
    current stream.  Returns (loss, alpha (H, W), d_image (H, W, C), d_lam,
    d_tgt)."""
    H, W = image.shape[:2]
    s = matting.assemble(image, lam, tgt, eps, _lib.ASSEMBLY_ACCURATE)
    alpha, _ = _solve(s, s.b, tol, max_iter, refine_mode)
    r = alpha - alpha_gt.reshape(-1).to(alpha.dtype)
    loss = torch.dot(r, r)
    mu, _ = _solve(s, 2.0 * r, tol, max_iter, refine_mode)
    gim, glam, gtg = matting.levin_backward(image, mu, alpha, lam, tgt, eps)
    gi = torch.zeros_like(image)
    gi[..., :3] = gim.to(image.dtype)
    return loss, alpha.view(H, W), gi, glam.view_as(lam).to(lam.dtype), gtg.view_as(tgt).to(tgt.dtype)
