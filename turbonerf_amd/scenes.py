"""Seeded synthetic scenes: Gaussian-blob density fields and sphere cameras.

A blob field is an (n, 8) float32 array, one row per blob:
``cx, cy, cz, s, p, r, g, b`` with the centre in world units. Density is
``sum_k p_k exp(-|x - c_k|^2 / (2 s_k^2))`` and colour is the
density-weighted mean of the blob colours. Scenes live in a box of side
``aabb`` centred at the origin, as the reference's BoundingBox(size) is
(src/models/bounding-box.cuh:20-23).
"""
from __future__ import annotations

import math

import numpy as np

from ._capi import BBox, Camera


def blob_field(seed: int, n_blobs: int = 8, aabb: float = 1.0, sigma_range=(0.03, 0.1), mu: float = 3.0,
               s: float = 0.5) -> np.ndarray:
    """Centres uniform in [0.2, 0.8]^3 of the box (world = (u - 0.5) * aabb),
    sigma uniform in sigma_range (scaled by aabb), peak lognormal(mu, s),
    rgb uniform in [0, 1]."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.2, 0.8, size=(n_blobs, 3))
    sig = rng.uniform(sigma_range[0], sigma_range[1], size=n_blobs)
    peak = rng.lognormal(mu, s, size=n_blobs)
    rgb = rng.uniform(0.0, 1.0, size=(n_blobs, 3))
    out = np.empty((n_blobs, 8), np.float32)
    out[:, 0:3] = (u - 0.5) * aabb
    out[:, 3] = sig * aabb
    out[:, 4] = peak
    out[:, 5:8] = rgb
    return out


def look_at_transform(eye: np.ndarray, target=(0.0, 0.0, 0.0)) -> np.ndarray:
    """3x4 camera-to-world matrix looking from `eye` at `target` (the
    origin by default). Columns: right, down, forward (the reference camera
    looks down its +z column, src/models/camera.cuh:119-132)."""
    eye = np.asarray(eye, float)
    fwd = np.asarray(target, float) - eye
    fwd /= np.linalg.norm(fwd)
    up = np.array([0.0, 0.0, 1.0])
    if abs(float(np.dot(fwd, up))) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    m = np.zeros((3, 4))
    m[:, 0] = right
    m[:, 1] = down
    m[:, 2] = fwd
    m[:, 3] = eye
    return m.astype(np.float32)


def make_camera(transform34: np.ndarray, width: int, height: int, fov_deg: float, near: float = 0.05,
                far: float = 100.0, dist_params=(0.0, 0.0, 0.0, 0.0, 0.0)) -> Camera:
    """Camera with the principal point at the image centre; fields as
    Camera(resolution, near, far, focal_length, principal_point, shift,
    transform, dist_params) builds them (src/models/camera.cuh:59-79).
    dist_params = (k1, k2, k3, p1, p2) of the lens model; all zero: pinhole."""
    f = 0.5 * width / math.tan(math.radians(fov_deg) * 0.5)
    cam = Camera()
    cam.resolution[0], cam.resolution[1] = width, height
    cam.near_plane, cam.far_plane = near, far
    cam.focal_length[0] = cam.focal_length[1] = np.float32(f)
    cam.resolution_f[0], cam.resolution_f[1] = float(width), float(height)
    cam.principal_point[0], cam.principal_point[1] = 0.5 * width, 0.5 * height
    cam.shift[0] = cam.shift[1] = 0.0
    # offset_px = shift * resolution - principal_point, in fp32 as the reference forms it
    cam.offset_px[0] = np.float32(np.float32(0.0) * np.float32(width) - np.float32(0.5 * width))
    cam.offset_px[1] = np.float32(np.float32(0.0) * np.float32(height) - np.float32(0.5 * height))
    for k, v in enumerate(np.asarray(transform34, np.float32).ravel()):
        cam.transform[k] = float(v)
    for k in range(5):
        cam.dist_params[k] = float(dist_params[k])
    cam.show_image_planes = 1
    return cam


def look_at_camera(eye, target, width: int, height: int, fov_deg: float, near: float = 0.05,
                   far: float = 100.0) -> Camera:
    """A pinhole camera at `eye` looking at `target`."""
    return make_camera(look_at_transform(eye, target), width, height, fov_deg, near, far)


def sphere_cameras(seed: int, n: int, radius: float, width: int, height: int, fov_deg: float,
                   near: float = 0.05, far: float = 100.0, dist_params=(0.0, 0.0, 0.0, 0.0, 0.0)):
    """n cameras at seeded uniform points of a sphere, looking at the origin."""
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    cams = (Camera * n)()
    for i in range(n):
        cams[i] = make_camera(look_at_transform(v[i] * radius), width, height, fov_deg, near, far, dist_params)
    return cams


def cameras_bytes(cams) -> np.ndarray:
    """Raw bytes of a ctypes Camera array (for upload to the device)."""
    import ctypes

    return np.frombuffer(ctypes.string_at(ctypes.addressof(cams), ctypes.sizeof(cams)), np.uint8).copy()


def bbox(aabb: float) -> BBox:
    return BBox.cube(aabb)


def struct_bytes(s) -> np.ndarray:
    import ctypes

    return np.frombuffer(ctypes.string_at(ctypes.addressof(s), ctypes.sizeof(s)), np.uint8).copy()
