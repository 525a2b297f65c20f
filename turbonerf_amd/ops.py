"""Thin torch-facing wrappers of the ``tn_*`` C ABI.

Every function takes CUDA (HIP) torch tensors, launches on torch's current
stream and returns new or updated tensors. These are the building blocks the
parity tests and the torch.autograd.Function wrappers (``autograd.py``) use;
the C++ Trainer/Renderer calls the same C ABI directly.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from ._capi import BBox, OccupancyGrid, call, ptr, query, stream_ptr

FULL = 0xFFFFFFFF


def _dev(t):
    if t is not None and not t.is_cuda:
        raise ValueError("expected a device tensor")
    return t


def upload_struct(s) -> torch.Tensor:
    raw = np.frombuffer(C.string_at(C.addressof(s), C.sizeof(s)), np.uint8).copy()
    return torch.from_numpy(raw).cuda()


def fill_uniform(n: int, seed: int, offset: int = 0, out: torch.Tensor | None = None) -> torch.Tensor:
    out = torch.empty(n, dtype=torch.float32, device="cuda") if out is None else out
    call("tn_fill_uniform", n, seed, offset, ptr(out), stream_ptr())
    return out


def grid_struct(bitfield: torch.Tensor, density: torch.Tensor | None, n_levels: int, res: int = 128) -> OccupancyGrid:
    """Occupancy-grid view over device tensors; the struct keeps the tensors
    alive (it holds raw pointers into them)."""
    g = OccupancyGrid(ptr(bitfield), ptr(density) if density is not None else None, n_levels, res)
    g._tensors = (bitfield, density)
    return g


@dataclass
class Rays:
    """Per-ray SoA buffers with component stride `stride`."""

    stride: int
    pix_rgba: torch.Tensor
    img_id: torch.Tensor
    ori: torch.Tensor
    dir: torch.Tensor
    t: torch.Tensor
    tmax: torch.Tensor
    alive: torch.Tensor
    n_steps: torch.Tensor
    offset: torch.Tensor

    @classmethod
    def empty(cls, stride: int) -> "Rays":
        f = lambda n: torch.zeros(n, dtype=torch.float32, device="cuda")
        u = lambda n: torch.zeros(n, dtype=torch.int32, device="cuda")
        return cls(stride, f(4 * stride), u(stride), f(3 * stride), f(3 * stride), f(stride), f(stride),
                   torch.zeros(stride, dtype=torch.uint8, device="cuda"), u(stride), u(stride))


@dataclass
class Samples:
    """Per-sample SoA buffers (stride = sample capacity)."""

    stride: int
    img_id: torch.Tensor
    t: torch.Tensor
    pos: torch.Tensor
    dir: torch.Tensor
    dt: torch.Tensor
    m_norm: torch.Tensor
    dt_norm: torch.Tensor

    @classmethod
    def empty(cls, stride: int) -> "Samples":
        f = lambda n: torch.zeros(n, dtype=torch.float32, device="cuda")
        return cls(stride, torch.zeros(stride, dtype=torch.int32, device="cuda"), f(stride), f(3 * stride),
                   f(3 * stride), f(stride), f(stride), f(stride))


def init_training_rays(rays: Rays, n_rays: int, n_images: int, img_w: int, img_h: int, n_rays_per_image: float,
                       bbox_dev: torch.Tensor, cams_dev: torch.Tensor, images_dev: torch.Tensor,
                       random: torch.Tensor, d_n_rays: torch.Tensor | None = None) -> None:
    call("tn_init_training_rays", n_rays, ptr(d_n_rays), rays.stride, n_images, img_w * img_h, img_w, img_h,
         n_rays_per_image, ptr(bbox_dev), ptr(cams_dev), ptr(images_dev), ptr(random), ptr(rays.pix_rgba),
         ptr(rays.img_id), ptr(rays.ori), ptr(rays.dir), ptr(rays.t), ptr(rays.tmax), ptr(rays.alive),
         stream_ptr())


def march_count(rays: Rays, n_rays: int, bbox: BBox, grid: OccupancyGrid, cone: float, dt_min: float,
                dt_max: float, max_steps: int = FULL, d_n_rays: torch.Tensor | None = None) -> None:
    call("tn_march_count", n_rays, ptr(d_n_rays), rays.stride, bbox, grid, cone, dt_min, dt_max, max_steps,
         ptr(rays.dir), ptr(rays.tmax), ptr(rays.alive), ptr(rays.ori), ptr(rays.t), ptr(rays.n_steps),
         stream_ptr())


def scan_select(rays: Rays, n_rays: int, max_samples: int, info: torch.Tensor, mode: int = 0,
                d_n_rays: torch.Tensor | None = None) -> None:
    nb = query("tn_scan_workspace_bytes", n_rays)
    ws = torch.empty(max(nb, 4), dtype=torch.uint8, device="cuda")
    call("tn_scan_select", mode, n_rays, ptr(d_n_rays), ptr(rays.n_steps), max_samples, ptr(rays.offset), ptr(info),
         ptr(ws), nb,
         stream_ptr())


def march_generate(rays: Rays, samples: Samples, n_rays_max: int, info: torch.Tensor, bbox: BBox, inv_aabb: float,
                   grid: OccupancyGrid, dt_min: float, dt_max: float, cone: float) -> None:
    if n_rays_max > rays.stride:
        raise ValueError("n_rays_max exceeds the ray capacity")
    call("tn_march_generate", n_rays_max, ptr(info), rays.stride, samples.stride, bbox, inv_aabb, grid, dt_min, dt_max, cone,
         ptr(rays.img_id), ptr(rays.ori), ptr(rays.dir), ptr(rays.t), ptr(rays.offset), ptr(rays.alive),
         ptr(rays.n_steps), ptr(samples.img_id), ptr(samples.t), ptr(samples.pos), ptr(samples.dir),
         ptr(samples.dt), ptr(samples.m_norm), ptr(samples.dt_norm), stream_ptr())


_DTYPE_CODE = {torch.float32: 0, torch.float16: 1}


def composite_train(info: torch.Tensor, n_rays_max: int, stride: int, sample_stride: int, steps: torch.Tensor,
                    offs: torch.Tensor, bg: torch.Tensor, target: torch.Tensor, density: torch.Tensor,
                    color: torch.Tensor, dt: torch.Tensor, grad_scale: float = 1.0, min_T: float = 1e-4,
                    fp64_accumulate: bool = False, scale_by_rays: bool = False):
    """Returns (loss[4*stride], dL_ddensity[sample_stride], dL_dcolor[3*sample_stride])."""
    code = 2 if fp64_accumulate else _DTYPE_CODE[density.dtype]
    if fp64_accumulate and density.dtype != torch.float32:
        raise ValueError("fp64 accumulation takes fp32 inputs")
    loss = torch.zeros(4 * stride, dtype=torch.float32, device="cuda")
    dLd = torch.zeros(sample_stride, dtype=torch.float32, device="cuda")
    dLc = torch.zeros(3 * sample_stride, dtype=density.dtype, device="cuda")
    nb = query("tn_composite_scratch_bytes", n_rays_max, sample_stride)
    scratch = torch.empty(nb // 4, dtype=torch.float32, device="cuda")
    call("tn_composite_train", code, ptr(info), n_rays_max, stride, sample_stride, grad_scale, int(scale_by_rays), min_T,
         ptr(steps), ptr(offs),
         ptr(bg), ptr(target), ptr(density), ptr(color), ptr(dt), ptr(loss), ptr(dLd), ptr(dLc), ptr(scratch), nb,
         stream_ptr())
    return loss, dLd, dLc


def composite_forward(info: torch.Tensor, stride: int, max_samples: int, n_rays_out: int, steps: torch.Tensor,
                      offs: torch.Tensor, density: torch.Tensor, color: torch.Tensor, dt: torch.Tensor,
                      raw_density: bool = True, min_T: float = 1e-4, fp64_accumulate: bool = False):
    code = 2 if fp64_accumulate else _DTYPE_CODE[density.dtype]
    rgba = torch.zeros(4 * n_rays_out, dtype=torch.float32, device="cuda")
    call("tn_composite_forward", code, ptr(info), stride, max_samples, n_rays_out, min_T, int(raw_density),
         ptr(steps), ptr(offs), ptr(density), ptr(color), ptr(dt), ptr(rgba), stream_ptr())
    return rgba


def reduce_loss(info: torch.Tensor, stride: int, loss: torch.Tensor, n_channels: int = 4) -> torch.Tensor:
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    call("tn_reduce_loss", ptr(info), stride, n_channels, ptr(loss), ptr(out), stream_ptr())
    return out


def blob_density_grid(grid: OccupancyGrid, blobs: torch.Tensor) -> None:
    call("tn_blob_density_grid", grid, ptr(blobs), blobs.numel() // 8, stream_ptr())


def blob_eval(info, n_max: int, stride: int, aabb: float, blobs: torch.Tensor, pos: torch.Tensor):
    sigma = torch.zeros(stride, dtype=torch.float32, device="cuda")
    rgb = torch.zeros(3 * stride, dtype=torch.float32, device="cuda")
    call("tn_blob_eval", ptr(info), n_max, stride, aabb, ptr(blobs), blobs.numel() // 8, ptr(pos), ptr(sigma),
         ptr(rgb), stream_ptr())
    return sigma, rgb


def occ_decay(grid: OccupancyGrid, factor: float) -> None:
    call("tn_occ_decay", grid, factor, stream_ptr())


def occ_update_bits(grid: OccupancyGrid, threshold: float) -> None:
    call("tn_occ_update_bits", grid, threshold, stream_ptr())


def occ_sample_points(n_cells: int, stride: int, start: int, grid: OccupancyGrid, level: int, inv_aabb: float,
                      random: torch.Tensor, pos: torch.Tensor) -> None:
    call("tn_occ_sample_points", n_cells, stride, start, grid, level, inv_aabb, ptr(random), ptr(pos), stream_ptr())


def occ_update_density(n_cells: int, start: int, grid: OccupancyGrid, level: int, sample_all: bool,
                       random: torch.Tensor, raw: torch.Tensor) -> None:
    code = _DTYPE_CODE[raw.dtype]
    call("tn_occ_update_density", code, n_cells, start, grid, level, int(sample_all), ptr(random), ptr(raw),
         stream_ptr())


def popcount(b: torch.Tensor) -> int:
    n = b.numel()
    nb = query("tn_popcount_workspace_bytes", n)
    ws = torch.empty(max(nb, 4), dtype=torch.uint8, device="cuda")
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    call("tn_popcount", ptr(b), n, ptr(out), ptr(ws), nb, stream_ptr())
    return int(out.item())


# ------------------------------------------------------------ hash grid / MLP
def hashgrid_layout(n_levels: int = 16, n_features: int = 2, log2_T: int = 19, base_res: int = 16,
                    per_level_scale: float | None = None, max_res: float = 2048.0) -> _capi.HashGrid:
    """Instant-NGP level table; per_level_scale defaults to
    exp((ln max_res - ln base_res) / (L - 1)) as the reference configures it
    (src/core/nerf-network.cu:86-91)."""
    import math

    if per_level_scale is None:
        per_level_scale = math.exp((math.log(max_res) - math.log(base_res)) / (n_levels - 1))
    g = _capi.HashGrid()
    call("tn_hashgrid_layout", n_levels, n_features, log2_T, base_res, per_level_scale, C.byref(g))
    return g


_PREC = {torch.float32: 0, torch.float16: 1}


def hashgrid_forward(grid: _capi.HashGrid, n: int, pos: torch.Tensor, stride: int, table: torch.Tensor,
                     feat_dtype=None, info: torch.Tensor | None = None, out_stride: int | None = None):
    """Features [32][out_stride]; fp32 unless feat_dtype says otherwise (the
    MLP kernels take fp32 features in both precisions)."""
    feat_dtype = torch.float32 if feat_dtype is None else feat_dtype
    out_stride = stride if out_stride is None else out_stride
    out = torch.zeros(2 * grid.n_levels * out_stride, dtype=feat_dtype, device="cuda")
    call("tn_hashgrid_forward", _PREC[table.dtype], _PREC[feat_dtype], n, ptr(info), stride, out_stride, grid,
         ptr(pos), ptr(table), ptr(out), stream_ptr())
    return out


def hashgrid_backward(grid: _capi.HashGrid, n: int, pos: torch.Tensor, stride: int, dfeat: torch.Tensor,
                      dtable: torch.Tensor, info: torch.Tensor | None = None, dfeat_stride: int | None = None,
                      binned: bool = False, planned: bool = False, split: int | None = None):
    """dtable += the table gradient. binned: the bin-sort kernels (with their
    workspace) instead of the atomic scatter-add; planned: the same in two
    calls, the plan from the positions first (tn_hashgrid_backward_plan),
    then the gradient passes; with `split` (a multiple of
    tn_hashgrid_backward_part_samples) the records of samples [split, n) are
    scattered first, then those of [0, split), then summed."""
    dfeat_stride = stride if dfeat_stride is None else dfeat_stride
    if planned:
        pb = query("tn_hashgrid_backward_plan_bytes", n, grid)
        wb = query("tn_hashgrid_backward_work_bytes", n, grid)
        if pb == 0 or wb == 0:
            raise ValueError("hash-grid layout exceeds the bin limits")
        plan = torch.empty(pb, dtype=torch.uint8, device="cuda")
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        call("tn_hashgrid_backward_plan", n, ptr(info), stride, grid, ptr(pos), ptr(plan), pb, stream_ptr())
        if split is None:
            call("tn_hashgrid_backward_planned", _PREC[dfeat.dtype], n, ptr(info), stride, dfeat_stride, grid,
                 ptr(pos), ptr(dfeat), ptr(dtable), ptr(plan), pb, ptr(work), wb, stream_ptr())
            return
        for lo, hi in ((split, n), (0, split)):
            call("tn_hashgrid_backward_scatter", _PREC[dfeat.dtype], n, ptr(info), stride, dfeat_stride, grid,
                 ptr(pos), ptr(dfeat), ptr(plan), pb, ptr(work), wb, lo, hi, stream_ptr())
        call("tn_hashgrid_backward_accumulate", n, grid, ptr(dtable), ptr(plan), pb, ptr(work), wb, stream_ptr())
        return
    if binned:
        nb = query("tn_hashgrid_backward_workspace_bytes", n, grid)
        if nb == 0:
            raise ValueError("hash-grid layout exceeds the bin limits")
        ws = torch.empty(nb, dtype=torch.uint8, device="cuda")
        call("tn_hashgrid_backward_binned", _PREC[dfeat.dtype], n, ptr(info), stride, dfeat_stride, grid, ptr(pos),
             ptr(dfeat), ptr(dtable), ptr(ws), nb, stream_ptr())
        return
    call("tn_hashgrid_backward", _PREC[dfeat.dtype], n, ptr(info), stride, dfeat_stride, grid, ptr(pos), ptr(dfeat),
         ptr(dtable), stream_ptr())


def appearance(bias: torch.Tensor | None, image_id: torch.Tensor | None = None, dz: torch.Tensor | None = None):
    """tn_appearance_t for the MLP calls (None without an embedding); bias [n][64] fp32."""
    if bias is None:
        return None
    if bias.dtype != torch.float32 or bias.numel() % 64:
        raise ValueError("appearance bias rows are fp32 [n][64]")
    return _capi.Appearance(ptr(bias), ptr(image_id), bias.numel() // 64, ptr(dz))


def _app_ref(app):
    return None if app is None else C.byref(app)


def mlp_forward(params: torch.Tensor, net_offset: int, n: int, feat: torch.Tensor, feat_stride: int,
                dir: torch.Tensor | None, stride: int, color: bool = True, info: torch.Tensor | None = None,
                app=None):
    """Density and colour; fp16 features (an fp16 network only) run the
    inference form, every layer's inputs rounded to fp16 once
    (tn_mlp_inference)."""
    prec = _PREC[params.dtype]
    f16 = feat.dtype == torch.float16
    if (feat.dtype != torch.float32 and not (f16 and prec == 1)) or feat.numel() < 32 * feat_stride:
        raise ValueError("the MLP kernels take fp32 features [32][feat_stride] (fp16 for an fp16 network)")
    dens = torch.zeros(stride, dtype=params.dtype, device="cuda")
    rgb = torch.zeros(3 * stride, dtype=params.dtype, device="cuda") if color else None
    call("tn_mlp_inference", prec, int(f16), n, ptr(info), stride, feat_stride, stride, ptr(params), net_offset,
         ptr(feat), ptr(dir), int(color), ptr(dens), ptr(rgb), _app_ref(app), stream_ptr())
    return dens, rgb


def appearance_bias(w3e: torch.Tensor, emb: torch.Tensor) -> torch.Tensor:
    """Per-image bias rows W3e emb[k]: w3e fp32 [64*16], emb fp32 [n*16] -> [n*64]."""
    n = emb.numel() // 16
    bias = torch.empty(64 * max(n, 1), dtype=torch.float32, device="cuda")
    call("tn_appearance_bias", n, ptr(w3e), ptr(emb), ptr(bias), stream_ptr())
    return bias


def appearance_backward(n: int, stride: int, image_id: torch.Tensor, dz: torch.Tensor, w3e: torch.Tensor,
                        emb: torch.Tensor, info: torch.Tensor | None = None):
    """(dL/dW3e [64*16], dL/demb [n_img*16]) from the colour layer's dz [64][stride]."""
    n_app = emb.numel() // 16
    g_w = torch.empty(64 * 16, dtype=torch.float32, device="cuda")
    g_e = torch.empty(n_app * 16, dtype=torch.float32, device="cuda")
    nb = query("tn_appearance_workspace_bytes", n_app)
    ws = torch.empty(max(nb, 4), dtype=torch.uint8, device="cuda")
    call("tn_appearance_backward", n, ptr(info), stride, ptr(image_id), ptr(dz), n_app, ptr(w3e), ptr(emb), ptr(g_w),
         ptr(g_e), ptr(ws), nb, stream_ptr())
    return g_w, g_e


def mlp_backward(params: torch.Tensor, net_offset: int, n: int, feat: torch.Tensor, feat_stride: int,
                 dir: torch.Tensor, stride: int, sample_t: torch.Tensor, dL_drgb: torch.Tensor, dL_ds: torch.Tensor,
                 loss_scale: float, grad_net: torch.Tensor, dL_ds_extra: torch.Tensor | None = None,
                 accumulate: bool = False, info: torch.Tensor | None = None, scale_by_rays: bool = False,
                 app=None):
    """Returns the fp32 feature gradients [32][feat_stride]; weight gradients go to grad_net."""
    prec = _PREC[params.dtype]
    if feat.dtype != torch.float32 or feat.numel() < 32 * feat_stride:
        raise ValueError("the MLP kernels take fp32 features [32][feat_stride]")
    if dL_drgb.dtype != params.dtype:
        raise ValueError("dL_drgb is in the network precision")
    dfeat = torch.zeros(32 * feat_stride, dtype=torch.float32, device="cuda")
    nb = query("tn_mlp_backward_workspace_bytes", n)
    ws = torch.empty(max(nb, 4), dtype=torch.uint8, device="cuda")
    call("tn_mlp_backward", prec, n, ptr(info), stride, feat_stride, ptr(params), net_offset, ptr(feat), ptr(dir),
         ptr(sample_t), ptr(dL_drgb), stride, ptr(dL_ds), ptr(dL_ds_extra), loss_scale, int(scale_by_rays), ptr(dfeat),
         ptr(grad_net), int(accumulate), ptr(ws), nb, _app_ref(app), stream_ptr())
    return dfeat


def adam_step(n_grid: int, n_net_end: int, loss_scale: float, lr: float, beta1: float, beta2: float, eps: float,
              l2: float, w: torch.Tensor, w_lp: torch.Tensor | None, grad: torch.Tensor, m: torch.Tensor,
              v: torch.Tensor, steps: torch.Tensor, scale_by_rays_info: torch.Tensor | None = None,
              d_step: torch.Tensor | None = None, lr_decay: float = 0.0) -> None:
    lp = 0 if w_lp is None else _PREC[w_lp.dtype]
    call("tn_adam_step", w.numel(), n_grid, n_net_end, loss_scale, ptr(scale_by_rays_info), lr, ptr(d_step), lr_decay,
         beta1, beta2, eps, l2, ptr(w), lp, ptr(w_lp), ptr(grad), ptr(m), ptr(v), ptr(steps), stream_ptr())
