"""TEST INFRASTRUCTURE ONLY: numpy/ctypes front end of the oracle library
(oracle/tn_oracle.cpp), the CPU restatement of the reference's hot-path math.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker. The product (turbonerf_amd/) never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "libtn_oracle.so"

_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32 = C.c_uint32
f32 = C.c_float
f64 = C.c_double
i32 = C.c_int


def _build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            _build()
        L = C.CDLL(str(LIB_PATH))
        sig = {
            "oc_uniform": (None, [C.c_uint64, C.c_uint64, C.c_uint64, f32p]),
            "oc_morton3d": (u32, [u32, u32, u32]),
            "oc_morton3d_invert": (u32, [u32]),
            "oc_fdiv_ru": (f32, [f32, f32]),
            "oc_grid_level_at": (i32, [i32, i32, f32, f32, f32, f32]),
            "oc_voxel_morton_index": (u32, [i32, i32, f32, f32, f32]),
            "oc_init_training_rays": (i32, [u32, u32, u32, u32, i32, i32, f32, C.c_void_p, C.c_void_p, u8p, f32p,
                                            f32p, u32p, f32p, f32p, f32p, f32p, u8p]),
            "oc_deactivate_rays": (None, [u32, u32, f32, f32, f32p, f32p, u8p]),
            "oc_march_count": (None, [u32, u32, C.c_void_p, u8p, i32, i32, f32, f32, f32, f32p, f32p, u8p, f32p,
                                      f32p, u32p, u32]),
            "oc_scan_select": (i32, [u32, u32p, u32, u32p, C.POINTER(u32), C.POINTER(u32)]),
            "oc_march_generate": (None, [u32, u32, u32, C.c_void_p, f32, u8p, i32, i32, f32, f32, f32, u32p, f32p, f32p,
                                         f32p, u32p, u8p, u32p, u32p, f32p, f32p, f32p, f32p, f32p, f32p]),
            "oc_composite_train_f32": (None, [u32, u32, u32, f32, f32, u32p, u32p, f32p, f32p, f32p, f32p, f32p, f32,
                                              f32p, f32p, f32p]),
            "oc_composite_train_f64": (None, [u32, u32, u32, f64, f64, u32p, u32p, f64p, f64p, f64p, f64p, f64p, f64,
                                              f64p, f64p, f64p]),
            "oc_distortion_f32": (None, [u32, f32, u32p, u32p, f32p, f32p, f32p, f32p, f32p]),
            "oc_distortion_f64": (None, [u32, f64, u32p, u32p, f64p, f64p, f64p, f64p, f64p]),
            "oc_distortion_bruteforce_f64": (f64, [u32, f64, f64p, f64p, f64p]),
            "oc_adam_step": (None, [u32, u32, u32, f32, f32, f32, f32, f32, f32, f32p, f32p, f32p, f32p, f32p,
                                    u32p]),
            "oc_occ_decay": (None, [u32, i32, f32, f32p]),
            "oc_occ_sample_points": (None, [u32, u32, u32, i32, i32, f32, f32p, f32p]),
            "oc_occ_update_density": (None, [u32, u32, u32, i32, i32, f32p, f32p, u8p, f32p]),
            "oc_occ_update_bits": (None, [u32, i32, f32, f32p, u8p]),
            "oc_popcount_bytes": (C.c_uint64, [C.c_uint64, u8p]),
            "oc_hashgrid_layout": (None, [i32, i32, i32, i32, f64, C.c_void_p]),
            "oc_hashgrid_fwd_f64": (None, [C.c_void_p, u32, u32, f32p, f64p, f64p]),
            "oc_hashgrid_bwd_f64": (None, [C.c_void_p, u32, u32, f32p, f64p, f64p]),
            "oc_sh4_f64": (None, [u32, u32, f32p, f64p]),
            "oc_hashgrid_fwd_f32": (None, [C.c_void_p, u32, u32, f32p, f32p, f32p]),
            "oc_hashgrid_bwd_f32": (None, [C.c_void_p, u32, u32, f32p, f32p, f32p]),
            "oc_mlp_train_f32": (None, [u32, u32, f32p, f32p, f32p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        f32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
            "oc_mlp_train_f64": (None, [u32, u32, f64p, f64p, f32p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        f64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
            "oc_nerf_fwd_f64": (None, [C.c_void_p, u32, u32, f32p, f32p, f64p, i32, f64p, f64p]),
            "oc_blob_eval": (None, [u32, f32p, u32, f32p, f32p, f32p]),
            "oc_composite_analytic_f64": (None, [u32, u32p, u32p, f32p, u32, f32, f32p, u32, f32p, f64, f64p]),
            "oc_lens_residual_jacobian": (None, [f32, f32, f32, f32, f32p, f32p]),
            "oc_lens_undistort": (None, [f32, f32, f32p, C.POINTER(f32), C.POINTER(f32)]),
            "oc_render_rays": (None, [u32, u32, u32, C.c_void_p, C.c_void_p, f32p, f32p, f32p, f32p, u8p]),
            "oc_render_fixed_step_analytic_f64": (None, [u32, u32, f32p, f32p, f32p, f32p, C.c_void_p, f32, u8p, i32, i32,
                                                         f32, f32, f32, u32, u32, f32p, f64, f64p]),
            "oc_num_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


# ------------------------------------------------------------------ wrappers
def uniform(n: int, seed: int, offset: int = 0) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().oc_uniform(n, seed, offset, out)
    return out


def morton3d(x, y, z) -> int:
    return int(lib().oc_morton3d(x, y, z))


def morton3d_invert(m) -> int:
    return int(lib().oc_morton3d_invert(m))


def fdiv_ru(a: float, b: float) -> float:
    return float(np.float32(lib().oc_fdiv_ru(a, b)))


def init_training_rays(n_rays, stride, n_images, img_w, img_h, n_rays_per_image, bbox, cameras, images, random):
    """Returns dict of SoA arrays (stride = `stride`). Buffers for dead rays keep
    their zero initial values, as the reference leaves them untouched."""
    pix = np.zeros(4 * stride, np.float32)
    img_id = np.zeros(stride, np.uint32)
    ori = np.zeros(3 * stride, np.float32)
    d = np.zeros(3 * stride, np.float32)
    t = np.zeros(stride, np.float32)
    tmax = np.zeros(stride, np.float32)
    alive = np.zeros(stride, np.uint8)
    rc = lib().oc_init_training_rays(n_rays, stride, n_images, img_w * img_h, img_w, img_h, n_rays_per_image,
                                     C.addressof(bbox), C.addressof(cameras[0]), c(images, np.uint8).ravel(),
                                     c(random, np.float32), pix, img_id, ori, d, t, tmax, alive)
    if rc != 0:
        raise RuntimeError(f"oc_init_training_rays: {rc}")
    return dict(pix_rgba=pix, img_id=img_id, ori=ori, dir=d, t=t, tmax=tmax, alive=alive)


def deactivate_rays(n_rays, stride, alpha_threshold, selection_probability, pix_rgba, random, alive):
    a = c(alive, np.uint8).copy()
    lib().oc_deactivate_rays(n_rays, stride, alpha_threshold, selection_probability, c(pix_rgba, np.float32),
                             c(random, np.float32), a)
    return a


def march_count(n_rays, stride, bbox, bitfield, n_levels, res, cone, dt_min, dt_max, rays, max_steps=0xFFFFFFFF):
    alive = rays["alive"].copy()
    n_steps = np.zeros(stride, np.uint32)
    lib().oc_march_count(n_rays, stride, C.addressof(bbox), c(bitfield, np.uint8), n_levels, res, cone, dt_min,
                         dt_max, rays["dir"], rays["tmax"], alive, rays["ori"], rays["t"], n_steps, max_steps)
    return n_steps, alive


def scan_select(n_steps, n_rays, max_samples):
    offs = np.zeros(len(n_steps), np.uint32)
    nr, ns = u32(), u32()
    rc = lib().oc_scan_select(n_rays, c(n_steps, np.uint32), max_samples, offs, C.byref(nr), C.byref(ns))
    return offs, int(nr.value), int(ns.value), int(rc)


def march_generate(n_rays, stride, sample_stride, bbox, inv_aabb, bitfield, n_levels, res, dt_min, dt_max, cone,
                   rays, offs, alive, n_steps):
    """`stride` is the ray buffers' stride, `sample_stride` the sample buffers'."""
    S = sample_stride
    o = dict(img_id=np.zeros(S, np.uint32), t=np.zeros(S, np.float32), pos=np.zeros(3 * S, np.float32),
             dir=np.zeros(3 * S, np.float32), dt=np.zeros(S, np.float32), m_norm=np.zeros(S, np.float32),
             dt_norm=np.zeros(S, np.float32))
    lib().oc_march_generate(n_rays, stride, S, C.addressof(bbox), inv_aabb, c(bitfield, np.uint8), n_levels, res,
                            dt_min, dt_max, cone, rays["img_id"], rays["ori"], rays["dir"], rays["t"],
                            c(offs, np.uint32), c(alive, np.uint8), c(n_steps, np.uint32), o["img_id"], o["t"],
                            o["pos"], o["dir"], o["dt"], o["m_norm"], o["dt_norm"])
    return o


def composite_train(n_rays, stride, sample_stride, steps, offs, bg, target, density, color, dt, grad_scale=1.0,
                    min_T=1e-4, fp64=True):
    """Per-ray arrays (bg, target, loss) use `stride`; per-sample arrays use
    `sample_stride`."""
    dt_ = np.float64 if fp64 else np.float32
    loss = np.zeros(4 * stride, dt_)
    dLd = np.zeros(sample_stride, dt_)
    dLc = np.zeros(3 * sample_stride, dt_)
    fn = lib().oc_composite_train_f64 if fp64 else lib().oc_composite_train_f32
    fn(n_rays, stride, sample_stride, 1.0 / n_rays, grad_scale, c(steps, np.uint32), c(offs, np.uint32), c(bg, dt_),
       c(target, dt_), c(density, dt_), c(color, dt_), c(dt, dt_), min_T, loss, dLd, dLc)
    return loss, dLd, dLc


def distortion(n_rays, steps, offs, density, m, dt, lam=0.01, fp64=True):
    dt_ = np.float64 if fp64 else np.float32
    n = len(density)
    loss = np.zeros(n_rays, dt_)
    grad = np.zeros(n, dt_)
    fn = lib().oc_distortion_f64 if fp64 else lib().oc_distortion_f32
    fn(n_rays, lam, c(steps, np.uint32), c(offs, np.uint32), c(density, dt_), c(m, dt_), c(dt, dt_), loss, grad)
    return loss, grad


def distortion_bruteforce(density, m, dt, lam=0.01) -> float:
    return float(lib().oc_distortion_bruteforce_f64(len(density), lam, c(density, np.float64), c(m, np.float64),
                                                    c(dt, np.float64)))


def adam_step(n_grid, n_net_end, loss_scale, lr, beta1, beta2, eps, l2, w, grad, m, v, steps):
    w = w.copy(); m = m.copy(); v = v.copy(); steps = steps.copy()
    wl = w.copy()
    lib().oc_adam_step(len(w), n_grid, n_net_end, loss_scale, lr, beta1, beta2, eps, l2, w, wl, c(grad, np.float32),
                       m, v, steps)
    return w, wl, m, v, steps


def occ_decay(density, n_levels, factor):
    d = density.copy()
    lib().oc_occ_decay(len(d) // n_levels, n_levels, factor, d)
    return d


def occ_sample_points(n_cells, stride, start, res, level, inv_aabb, random):
    pos = np.zeros(3 * stride, np.float32)
    lib().oc_occ_sample_points(n_cells, stride, start, res, level, inv_aabb, c(random, np.float32), pos)
    return pos


def occ_update_density(n_cells, start, volume, level, sample_all, random, raw, bits, density):
    d = density.copy()
    lib().oc_occ_update_density(n_cells, start, volume, level, int(sample_all), c(random, np.float32),
                                c(raw, np.float32), c(bits, np.uint8), d)
    return d


def occ_update_bits(density, volume, n_levels, thr):
    bits = np.zeros(volume, np.uint8)
    lib().oc_occ_update_bits(volume, n_levels, thr, c(density, np.float32), bits)
    return bits


def popcount(b) -> int:
    return int(lib().oc_popcount_bytes(len(b), c(b, np.uint8)))


def blob_eval(xyz, blobs):
    xyz = c(xyz, np.float32).reshape(-1, 3)
    n = len(xyz)
    s = np.zeros(n, np.float32)
    rgb = np.zeros(3 * n, np.float32)
    lib().oc_blob_eval(n, xyz.ravel(), len(blobs), c(blobs, np.float32).ravel(), s, rgb)
    return s, rgb.reshape(n, 3)


def composite_analytic(n_rays, steps, offs, pos, stride, aabb, dt, blobs, min_T=1e-4):
    out = np.zeros(4 * n_rays, np.float64)
    lib().oc_composite_analytic_f64(n_rays, c(steps, np.uint32), c(offs, np.uint32), c(pos, np.float32), stride,
                                    aabb, c(dt, np.float32), len(blobs), c(blobs, np.float32).ravel(), min_T, out)
    return out.reshape(4, n_rays)


def mlp_train(net, feat, dir01, G, dL_ds, t, loss_scale, extra=None, fp64=True, dW=None):
    """MLP forward (+ backward when G is given) over n samples. feat (32, n),
    dir01 (3, n), G = dL/drgb (3, n) or None, dL_ds / t / extra (n,). Returns
    (density s0 (n,), rgb (3, n), dfeat (32, n), dW (9408,)); dW accumulates
    into the given array."""
    dt_ = np.float64 if fp64 else np.float32
    n = np.shape(feat)[1]
    dens = np.zeros(n, dt_)
    rgb = np.zeros((3, n), dt_)
    dfeat = np.zeros((32, n), dt_)
    if dW is None:
        dW = np.zeros(9408, dt_)
    keep = []

    def p(a, t_):
        if a is None:
            return None
        a = c(a, t_)
        keep.append(a)
        return a.ctypes.data

    fn = lib().oc_mlp_train_f64 if fp64 else lib().oc_mlp_train_f32
    fn(n, n, c(net, dt_), c(feat, dt_), c(dir01, np.float32), p(G, dt_), p(dL_ds, np.float32), p(extra, np.float32),
       p(t, np.float32), loss_scale, dens.ctypes.data, rgb.ctypes.data, dfeat.ctypes.data, dW.ctypes.data)
    return dens, rgb, dfeat, dW


def lens_residual_jacobian(x, y, xd, yd, k):
    """(fx, fy, fx_x, fx_y, fy_x, fy_y) of the lens model at (x, y) as float32."""
    out = np.zeros(6, np.float32)
    lib().oc_lens_residual_jacobian(x, y, xd, yd, np.asarray(k, np.float32), out)
    return out


def lens_undistort(xd, yd, k):
    """Undistorted normalised point of the distorted (xd, yd), float32 pair."""
    xu, yu = f32(), f32()
    lib().oc_lens_undistort(xd, yd, np.asarray(k, np.float32), C.byref(xu), C.byref(yu))
    return np.float32(xu.value), np.float32(yu.value)


def render_rays(n, start, stride, camera, bbox):
    """Render rays of pixels [start, start + n): dict of SoA arrays (stride)."""
    o = dict(ori=np.zeros(3 * stride, np.float32), dir=np.zeros(3 * stride, np.float32),
             t=np.zeros(stride, np.float32), tmax=np.zeros(stride, np.float32), alive=np.zeros(stride, np.uint8))
    lib().oc_render_rays(n, start, stride, C.addressof(camera), C.addressof(bbox), o["ori"], o["dir"], o["t"],
                         o["tmax"], o["alive"])
    return o


def render_fixed_step_analytic(n_rays, stride, ori, dir, t, tmax, bbox, bitfield, n_levels, res, cone, dt_min, dt_max,
                               max_steps, blobs, min_T=1e-4, inv_aabb=1.0):
    """(4, n_rays) rgba of the reference's fixed-step render loop over the blob
    field; inv_aabb = 1 / the scene size (the network step's normalisation)."""
    out = np.zeros(4 * n_rays, np.float64)
    lib().oc_render_fixed_step_analytic_f64(n_rays, stride, c(ori, np.float32), c(dir, np.float32), c(t, np.float32),
                                            c(tmax, np.float32), C.addressof(bbox), inv_aabb, c(bitfield, np.uint8), n_levels,
                                            res, cone, dt_min, dt_max, max_steps, len(blobs),
                                            c(blobs, np.float32).ravel(), min_T, out)
    return out.reshape(4, n_rays)


def num_threads() -> int:
    return int(lib().oc_num_threads())
