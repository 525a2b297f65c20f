"""Synthetic training data: render a Gaussian-blob field (scenes.blob_field)
to RGBA8 images on the GPU, with the same marching, compositing and ray
kernels the renderer uses (no network: density and colour are evaluated
analytically at every sample).

Pixels store straight (un-premultiplied) colour and alpha, the layout
training images have; training composites them over a random background
(reference src/utils/training-kernels-fused.cuh:443-446).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import ops, scenes
from ._capi import call, ptr, stream_ptr

DT_MIN = math.sqrt(3.0) / 1024.0
CONE = 1.0 / 256.0


def blob_occupancy(blobs: torch.Tensor, n_levels: int, threshold: float = 0.01, res: int = 128):
    """Bitfield and per-cascade densities of the blob field at cell centres."""
    vol = res ** 3
    bits = torch.zeros(vol, dtype=torch.uint8, device="cuda")
    dens = torch.zeros(vol * n_levels, dtype=torch.float32, device="cuda")
    grid = ops.grid_struct(bits, dens, n_levels, res)
    ops.blob_density_grid(grid, blobs)
    ops.occ_update_bits(grid, threshold)
    return bits, dens, grid


def render_blob_images(blobs_np: np.ndarray, cams, width: int, height: int, aabb: float = 1.0,
                       max_steps: int = 1024, rays_per_pass: int = 1 << 18, samples_per_pass: int = 1 << 23,
                       threshold: float = 0.01) -> np.ndarray:
    n_levels = int(math.log2(aabb)) + 1
    blobs = torch.from_numpy(np.ascontiguousarray(blobs_np, np.float32).ravel()).cuda()
    bits, dens, grid = blob_occupancy(blobs, n_levels, threshold)
    bbox = scenes.bbox(aabb)
    R, S = rays_per_pass, samples_per_pass
    rays = ops.Rays.empty(R)
    pixel = torch.zeros(R, dtype=torch.int32, device="cuda")
    samples = ops.Samples.empty(S)
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    n_pix = width * height
    out = np.zeros((len(cams), height, width, 4), np.uint8)
    image = torch.zeros(n_pix * 4, dtype=torch.float32, device="cuda")
    dt_max = aabb * DT_MIN
    for c in range(len(cams)):
        image.zero_()
        p = 0
        while p < n_pix:
            m = min(R, n_pix - p)
            call("tn_render_rays", m, p, R, cams[c], bbox, None, ptr(rays.ori), ptr(rays.dir), ptr(rays.t),
                 ptr(rays.tmax), ptr(pixel), ptr(rays.alive), stream_ptr())
            ops.march_count(rays, m, bbox, grid, CONE, DT_MIN, dt_max, max_steps)
            ops.scan_select(rays, m, S, info, mode=1)
            ops.march_generate(rays, samples, m, info, bbox, 1.0 / aabb, grid, DT_MIN, dt_max, CONE)
            sigma, rgb = ops.blob_eval(info, S, S, aabb, blobs, samples.pos)
            rgba = ops.composite_forward(info, S, S, R, rays.n_steps, rays.offset, sigma, rgb, samples.dt,
                                         raw_density=False)
            call("tn_render_write_rgba", m, ptr(info), R, ptr(pixel), ptr(rays.alive), ptr(rgba), ptr(image),
                 stream_ptr())
            kept = int(info[0].item())
            if kept == 0:
                raise RuntimeError("a ray exceeds the sample budget")
            p += kept
        img = image.view(height, width, 4)
        a = img[..., 3:4].clamp(0, 1)
        rgb = torch.where(a > 1e-6, img[..., :3] / a.clamp_min(1e-6), torch.zeros_like(img[..., :3])).clamp(0, 1)
        out[c] = torch.cat([rgb, a], -1).mul(255.0).round().to(torch.uint8).cpu().numpy()
    return out


def blob_dataset(seed: int, n_cams: int, width: int, height: int, fov_deg: float = 50.0, radius: float = 2.0,
                 aabb: float = 1.0, n_blobs: int = 8, dist_params=(0.0, 0.0, 0.0, 0.0, 0.0)):
    """(blobs, ctypes camera array, images) for a seeded blob scene; the
    images are rendered through the cameras' lens model (dist_params)."""
    blobs = scenes.blob_field(seed, n_blobs, aabb)
    cams = scenes.sphere_cameras(seed + 1, n_cams, radius * aabb, width, height, fov_deg, dist_params=dist_params)
    images = render_blob_images(blobs, cams, width, height, aabb)
    return blobs, cams, images


def to_pyturbonerf_cameras(tn, cams):
    """ctypes cameras -> PyTurboNeRF.Camera objects."""
    out = []
    for c in cams:
        t = np.array(list(c.transform), np.float32).reshape(3, 4)
        out.append(tn.Camera((c.resolution[0], c.resolution[1]), c.near_plane, c.far_plane,
                             (c.focal_length[0], c.focal_length[1]), (c.principal_point[0], c.principal_point[1]),
                             (c.shift[0], c.shift[1]), tn.Transform4f(t), tn.DistortionParams(*c.dist_params)))
    return out
