"""torch.autograd.Function wrappers over the hand-written adjoint kernels.

* ``HashGridEncoding``: hash-grid forward; backward is the scatter-add into the
  table gradient (the exact adjoint of the forward, which is linear in the
  table). Positions get no gradient, as in the reference.
* ``NerfMLP``: fused density + colour MLP forward; backward through the fused
  MFMA backward kernel with unit loss scale and no distance scaling, i.e. the
  plain vector-Jacobian product.
* ``CompositeLoss``: volume-rendering compositing over a random background +
  smooth-L1 against the targets, summed over the rgb channels and averaged
  over rays (reference src/utils/training-kernels-fused.cuh:381-498). The
  kernel recomputes transmittance in its backward sweep instead of storing it;
  the gradients come out of the same launch and are scaled by the incoming
  gradient here.
* ``DistortionLoss``: mip-NeRF 360 distortion loss summed over rays, O(N)
  prefix-sum form (reference training-kernels-fused.cuh:13-224).

Tensors are SoA and feature-major as the kernels take them: features
``(32, n)``, directions/colours ``(3, n)``, per-ray ``bg (3, n_rays)`` and
``target (4, n_rays)``. Everything runs on torch's current stream.
"""
from __future__ import annotations

import torch

from . import ops
from ._capi import call, ptr, query, stream_ptr

__all__ = ["HashGridEncoding", "NerfMLP", "CompositeLoss", "DistortionLoss"]


def _contig(t: torch.Tensor) -> torch.Tensor:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    return t.contiguous()


class HashGridEncoding(torch.autograd.Function):
    """features (2L, n) = encode(pos (3, n) in [0,1], table (P,))."""

    @staticmethod
    def forward(ctx, table: torch.Tensor, pos: torch.Tensor, grid):
        table, pos = _contig(table), _contig(pos)
        n = pos.shape[1]
        feat = ops.hashgrid_forward(grid, n, pos, n, table)
        ctx.save_for_backward(pos)
        ctx.grid = grid
        ctx.table_dtype = table.dtype
        ctx.n_table = table.numel()
        return feat.view(2 * grid.n_levels, n)

    @staticmethod
    def backward(ctx, grad_feat: torch.Tensor):
        (pos,) = ctx.saved_tensors
        n = pos.shape[1]
        dtable = torch.zeros(ctx.n_table, dtype=torch.float32, device=pos.device)
        g = _contig(grad_feat)
        if g.dtype not in (torch.float32, torch.float16):
            g = g.float()
        ops.hashgrid_backward(ctx.grid, n, pos, n, g, dtable)
        return dtable.to(ctx.table_dtype), None, None


class NerfMLP(torch.autograd.Function):
    """(density s (n,), rgb (3, n)) = MLP(net (9408,), features (32, n), dir01 (3, n))."""

    @staticmethod
    def forward(ctx, net: torch.Tensor, feat: torch.Tensor, dir01: torch.Tensor):
        net, feat, dir01 = _contig(net), _contig(feat.float()), _contig(dir01.float())
        n = feat.shape[1]
        dens, rgb = ops.mlp_forward(net, 0, n, feat, n, dir01, n, color=True)
        ctx.save_for_backward(net, feat, dir01)
        return dens, rgb.view(3, n)

    @staticmethod
    def backward(ctx, grad_dens, grad_rgb):
        net, feat, dir01 = ctx.saved_tensors
        n = feat.shape[1]
        if grad_dens is None:
            grad_dens = torch.zeros(n, dtype=torch.float32, device=net.device)
        if grad_rgb is None:
            grad_rgb = torch.zeros(3, n, dtype=net.dtype, device=net.device)
        ones = torch.ones(n, dtype=torch.float32, device=net.device)  # t >= 1: no distance scaling
        grad_net = torch.zeros(net.numel(), dtype=torch.float32, device=net.device)
        dfeat = ops.mlp_backward(net, 0, n, feat, n, dir01, n, ones, _contig(grad_rgb.to(net.dtype)),
                                 _contig(grad_dens.float()), 1.0, grad_net)
        return grad_net.to(net.dtype), dfeat.view(32, n), None


class CompositeLoss(torch.autograd.Function):
    """Scalar (1/n_rays) * sum over rays and rgb of smoothL1(composite - target)."""

    @staticmethod
    def forward(ctx, density, color, dt, steps, offsets, bg, target, min_T: float = 1e-4):
        density, color, dt = _contig(density), _contig(color), _contig(dt)
        steps, offsets, bg, target = _contig(steps), _contig(offsets), _contig(bg), _contig(target)
        n_rays = steps.numel()
        S = density.numel()
        info = torch.tensor([n_rays, S, 0, 0], dtype=torch.int32, device=density.device)
        loss, dLd, dLc = ops.composite_train(info, n_rays, n_rays, S, steps, offsets, bg, target, density,
                                             color, dt, grad_scale=1.0, min_T=min_T)
        ctx.save_for_backward(dLd, dLc)
        ctx.shapes = (density.dtype, color.shape)
        return loss.view(4, n_rays)[:3].sum() / n_rays

    @staticmethod
    def backward(ctx, grad):
        dLd, dLc = ctx.saved_tensors
        dtype, cshape = ctx.shapes
        return (grad * dLd).to(dtype), (grad * dLc.float()).to(dtype).view(cshape), None, None, None, None, None, None


class DistortionLoss(torch.autograd.Function):
    """Scalar sum over rays of the distortion loss (weight `lam`)."""

    @staticmethod
    def forward(ctx, density, m_norm, dt_norm, steps, offsets, lam: float = 0.01):
        density, m_norm, dt_norm = _contig(density), _contig(m_norm), _contig(dt_norm)
        steps, offsets = _contig(steps), _contig(offsets)
        n_rays = steps.numel()
        S = density.numel()
        code = {torch.float32: 0, torch.float16: 1}[density.dtype]
        info = torch.tensor([n_rays, S, 0, 0], dtype=torch.int32, device=density.device)
        loss = torch.zeros(n_rays, dtype=torch.float32, device=density.device)
        grad = torch.zeros(S, dtype=torch.float32, device=density.device)
        nb = query("tn_distortion_scratch_bytes", n_rays, S)
        scratch = torch.empty(max(nb, 4), dtype=torch.uint8, device=density.device)
        call("tn_distortion", code, ptr(info), n_rays, S, lam, ptr(steps), ptr(offsets), ptr(density), ptr(m_norm),
             ptr(dt_norm), ptr(loss), ptr(grad), ptr(scratch), nb, stream_ptr())
        ctx.save_for_backward(grad)
        ctx.dtype = density.dtype
        return loss.sum()

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (g * grad).to(ctx.dtype), None, None, None, None, None
