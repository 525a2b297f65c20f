"""Several NeRFs in one render, and spatial effects (reference
src/core/renderer.cu:344-521, src/utils/rendering-kernels.cu:175-606,
src/effects/spatial/repeater.cu:8-76), on synthetic blob-field NeRFs.

* One NeRF through the shared fixed-step loop against the oracle's
  restatement of the reference's render loop (oracle
  render_fixed_step_analytic): the same sample positions (float32 steps,
  contraction off), so only fp32 vs fp64 evaluation and compositing differ.
* Two overlapping NeRFs (one moved by a rigid transform) against the
  reference's blend restated here in numpy: per shared world sample each
  NeRF inside its box on an occupied cell adds alpha = 1 - exp(-sigma dt)
  and alpha * colour; the sample composites with the sum capped at 1 and the
  alpha-weighted mean colour; the rays come from the GPU's ray kernel.
* Two NeRFs far apart: each pixel equals the render of the one NeRF its ray
  meets, bit for bit.
* A repeater (source = the NeRF's box, extended over two more boxes along x)
  renders as the NeRF plus two copies moved by one and two box widths.
"""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT_MIN = math.sqrt(3.0) / 1024.0
CONE = 1.0 / 256.0


def _nerf(tn, mgr, blobs, aabb, transform=None):
    nerf = mgr.create()
    nerf.set_analytic_field(blobs, aabb, 0.01)
    if transform is not None:
        nerf.transform = tn.Transform4f(np.asarray(transform, np.float32))
    return nerf


def _render(tn, synthetic, cam, renderables, shared=False):
    buf = tn.CPURenderBuffer()
    r = tn.Renderer()
    r._force_shared_steps = shared
    r.submit(tn.RenderRequest(synthetic.to_pyturbonerf_cameras(tn, [cam])[0], renderables, buf))
    return np.array(buf.get_rgba())


def test_shared_steps_one_nerf_vs_reference_loop(cuda):
    import oracle
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    aabb, levels, seed = 16.0, 5, 404
    W, H = 120, 68
    blobs = scenes.blob_field(seed, 256, aabb, sigma_range=(0.004, 0.012))
    mgr = tn.NeRFManager()
    nerf = _nerf(tn, mgr, blobs, aabb)
    bits = nerf.get_occupancy_bitfield()
    cam = scenes.sphere_cameras(seed + 1, 1, 24.0, W, H, 50.0)[0]
    got = _render(tn, synthetic, cam, [tn.Renderable(nerf)], shared=True).reshape(-1, 4).T
    n = W * H
    rr = oracle.render_rays(n, 0, n, cam, scenes.bbox(aabb))
    tmax = np.where(rr["alive"] != 0, rr["tmax"], np.float32(-1)).astype(np.float32)
    dt = float(np.float32(DT_MIN))
    ref = oracle.render_fixed_step_analytic(n, n, rr["ori"], rr["dir"], rr["t"], tmax, scenes.bbox(aabb), bits, levels,
                                            128, CONE, dt, 8 * dt, 1 << 30, blobs, inv_aabb=1.0 / aabb)
    err = np.abs(got - ref)
    print(f"[render multi] one NeRF, shared steps: alpha mean {ref[3].mean():.3f}; |err| max {err.max():.2e} "
          f"p99.9 {np.percentile(err, 99.9):.2e} mean {err.mean():.2e}", flush=True)
    assert ref[3].max() > 0.5 and ref[3].mean() > 0.02
    # measured: max 9e-5, mean 1.2e-7
    assert err.mean() < 2e-6 and err.max() < 1e-3
    mgr.destroy(nerf)


# ------------------------------------------------------------ restated blend
def _morton(x, y, z):
    def spread(v):
        v = v.astype(np.uint64) & 0x3FF
        v = (v | (v << 16)) & 0x030000FF
        v = (v | (v << 8)) & 0x0300F00F
        v = (v | (v << 4)) & 0x030C30C3
        v = (v | (v << 2)) & 0x09249249
        return v
    return (spread(x) | (spread(y) << 1) | (spread(z) << 2)).astype(np.int64)


def _occupied(bits, levels, res, x, y, z, dt):
    """Reference OccupancyGrid::get_grid_level_at and is_occupied_at
    (occupancy-grid.cuh:185-260), float32, vectorised."""
    lvl = np.full(x.shape, levels - 1)
    found = np.zeros(x.shape, bool)
    for k in range(levels):
        half, cell = np.float32(0.5 * 2 ** k), np.float32(2 ** k / res)
        hit = ~found & (np.abs(x) < half) & (np.abs(y) < half) & (np.abs(z) < half) & (cell > dt)
        lvl[hit] = k
        found |= hit
    s = (np.float32(1.0) / np.float32(2.0) ** lvl.astype(np.float32)).astype(np.float32)
    idx = [np.minimum(np.maximum((v * s + np.float32(0.5)) * np.float32(res), 0).astype(np.int64), res - 1)
           for v in (x, y, z)]
    return ((bits[_morton(*idx)] >> lvl) & 1).astype(bool)


def _blend_reference(rays, nerfs, n_steps_max=1 << 16, min_T=1e-4):
    """The reference's shared-sample blend (rendering-kernels.cu:264-606),
    rays from the GPU's ray kernel: rgba [4][n] in float64."""
    import oracle

    ori, dir_, tmax = rays["ori"], rays["dir"], rays["tmax"]
    n = ori.shape[1]
    t = np.zeros(n, np.float32)
    alive = rays["alive"].astype(bool).copy()
    T = np.ones(n)
    out = np.zeros((4, n))
    dt_min = np.float32(DT_MIN)
    for _ in range(n_steps_max):
        live = alive & (t <= tmax)
        if not live.any():
            break
        alive &= live
        p = [ori[c] + t * dir_[c] for c in range(3)]  # float32, no contraction
        dt = np.minimum(np.maximum(t * np.float32(CONE), dt_min), np.float32(8) * dt_min).astype(np.float32)
        sa = np.zeros(n)
        sc = np.zeros((3, n))
        cnt = np.zeros(n, int)
        for nf in nerfs:
            M = nf["to_local"]
            lx, ly, lz = (M[r, 0] * p[0] + M[r, 1] * p[1] + M[r, 2] * p[2] + M[r, 3] for r in range(3))
            b = nf["box"]
            inside = live & (lx >= b[0]) & (lx <= b[3]) & (ly >= b[1]) & (ly <= b[4]) & (lz >= b[2]) & (lz <= b[5])
            dl = [M[r, 0] * dir_[0] + M[r, 1] * dir_[1] + M[r, 2] * dir_[2] for r in range(3)]
            scale = np.sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]).astype(np.float32)
            ndt = (np.float32(1.0 / nf["aabb"]) * (scale * dt)).astype(np.float32)
            ok = inside.copy()
            ok[inside] = _occupied(nf["bits"], nf["levels"], 128, lx[inside], ly[inside], lz[inside], ndt[inside])
            if not ok.any():
                continue
            sig, rgb = oracle.blob_eval(np.stack([lx[ok], ly[ok], lz[ok]], 1), nf["blobs"])
            a = 1.0 - np.exp(-sig.astype(np.float64) * ndt[ok])
            sa[ok] += a
            sc[:, ok] += rgb.T * a
            cnt[ok] += 1
        m = cnt > 0
        inv = np.where(sa > 1e-4, 1.0 / np.maximum(sa, 1e-300), 0.0)
        a = np.minimum(sa, 1.0)
        w = np.where(m, a * T, 0.0)
        out[:3] += w * sc * inv
        out[3] += w
        T = np.where(m, T * (1.0 - a), T)
        alive &= T >= min_T
        t = np.where(live, t + dt, t).astype(np.float32)
    return out


def _rays_multi(tn, cam, nerfs_c, n):
    import torch

    from turbonerf_amd import _capi

    arr = (_capi.RenderNerf * len(nerfs_c))(*nerfs_c)
    d_nerfs = torch.from_numpy(np.frombuffer(bytes(arr), np.uint8).copy()).cuda()
    f = lambda k=1: torch.zeros(k * n, dtype=torch.float32, device="cuda")
    ori, dir_, t, tmax, trans = f(3), f(3), f(), f(), f()
    steps = torch.zeros(n, dtype=torch.int32, device="cuda")
    pixel = torch.zeros(n, dtype=torch.int32, device="cuda")
    alive = torch.zeros(n, dtype=torch.uint8, device="cuda")
    P = _capi.ptr
    _capi.call("tn_render_rays_multi", n, 0, n, cam, len(nerfs_c), P(d_nerfs), P(ori), P(dir_), P(t), P(tmax),
               P(steps), P(pixel), P(alive), P(trans), _capi.stream_ptr())
    torch.cuda.synchronize()
    return dict(ori=ori.cpu().numpy().reshape(3, n), dir=dir_.cpu().numpy().reshape(3, n),
                tmax=tmax.cpu().numpy(), alive=alive.cpu().numpy())


def _placement(M, box, aabb):
    from turbonerf_amd import _capi

    nf = _capi.RenderNerf()
    for i, v in enumerate(np.asarray(M, np.float32).ravel()):
        nf.to_local[i] = float(v)
    nf.render_box = _capi.BBox(*[float(v) for v in box])
    nf.repeat = 0
    nf.inv_aabb = 1.0 / aabb
    nf.training_min[0] = nf.training_min[1] = nf.training_min[2] = -0.5 * aabb
    nf.training_size = aabb
    return nf


def _rigid(angle, axis, shift):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K
    return np.concatenate([R, np.asarray(shift, float)[:, None]], 1)


def _inverse(T):
    R, t = T[:, :3], T[:, 3]
    return np.concatenate([R.T, (-R.T @ t)[:, None]], 1)


def test_two_overlapping_nerfs_vs_restated_blend(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    W, H = 64, 48
    blobs_a = scenes.blob_field(71, 8, 1.0)
    blobs_b = scenes.blob_field(72, 8, 1.0)
    Tb = _rigid(0.5, [0.2, 1.0, 0.3], [0.3, 0.05, -0.1])
    mgr = tn.NeRFManager()
    na = _nerf(tn, mgr, blobs_a, 1.0)
    nb = _nerf(tn, mgr, blobs_b, 1.0, Tb)
    cam = scenes.sphere_cameras(73, 1, 2.2, W, H, 50.0)[0]
    got = _render(tn, synthetic, cam, [tn.Renderable(na), tn.Renderable(nb)]).reshape(-1, 4).T
    box = (-0.5, -0.5, -0.5, 0.5, 0.5, 0.5)
    # the transform the renderer applies: the inverse of the placement, as float32 rows
    Mb = np.asarray(tn.Transform4f(np.asarray(Tb, np.float32)).inverse().to_matrix(), np.float32)[:3]
    Ma = np.eye(4, dtype=np.float32)[:3]
    rays = _rays_multi(tn, cam, [_placement(Ma, box, 1.0), _placement(Mb, box, 1.0)], W * H)
    nerfs = [dict(to_local=Ma, box=box, aabb=1.0, bits=na.get_occupancy_bitfield(), levels=1, blobs=blobs_a),
             dict(to_local=Mb, box=box, aabb=1.0, bits=nb.get_occupancy_bitfield(), levels=1, blobs=blobs_b)]
    ref = _blend_reference(rays, nerfs)
    err = np.abs(got - ref)
    both = (ref[3] > 0.05)
    print(f"[render multi] two NeRFs: alpha mean {ref[3].mean():.3f}, {both.mean():.2f} of pixels covered; |err| max "
          f"{err.max():.2e} p99.9 {np.percentile(err, 99.9):.2e} mean {err.mean():.2e}", flush=True)
    assert ref[3].mean() > 0.02
    # measured: max 1e-4, mean 3e-7
    assert err.mean() < 3e-6 and err.max() < 1e-3
    for x in (na, nb):
        mgr.destroy(x)


def test_disjoint_nerfs_equal_their_single_renders(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    W, H = 80, 40
    mgr = tn.NeRFManager()
    left = np.concatenate([np.eye(3), [[-1.2], [0.0], [0.0]]], 1)
    right = np.concatenate([np.eye(3), [[1.2], [0.0], [0.0]]], 1)
    na = _nerf(tn, mgr, scenes.blob_field(81, 8, 1.0), 1.0, left)
    nb = _nerf(tn, mgr, scenes.blob_field(82, 8, 1.0), 1.0, right)
    # a camera on the +z axis looking at the origin: no ray meets both boxes
    cam = scenes.look_at_camera([0.0, 0.0, 4.0], [0.0, 0.0, 0.0], W, H, 60.0)
    both = _render(tn, synthetic, cam, [tn.Renderable(na), tn.Renderable(nb)])
    a = _render(tn, synthetic, cam, [tn.Renderable(na)], shared=True)
    b = _render(tn, synthetic, cam, [tn.Renderable(nb)], shared=True)
    assert a[..., 3].max() > 0.3 and b[..., 3].max() > 0.3
    assert not ((a[..., 3] > 0) & (b[..., 3] > 0)).any()
    assert np.array_equal(both, a + b)
    for x in (na, nb):
        mgr.destroy(x)


def test_repeater_equals_moved_copies(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    W, H = 96, 40
    blobs = scenes.blob_field(91, 8, 1.0)
    mgr = tn.NeRFManager()
    base = _nerf(tn, mgr, blobs, 1.0)
    copies = [_nerf(tn, mgr, blobs, 1.0, np.concatenate([np.eye(3), [[dx], [0.0], [0.0]]], 1)) for dx in (1.0, 2.0)]
    cam = scenes.look_at_camera([1.0, 0.3, 4.5], [1.0, 0.0, 0.0], W, H, 55.0)
    rep = tn.RepeaterEffect(tn.BoundingBox(1.0), tn.BoundingBox(-0.5, -0.5, -0.5, 2.5, 0.5, 0.5))
    repeated = _render(tn, synthetic, cam, [tn.Renderable(base, [rep])])
    three = _render(tn, synthetic, cam, [tn.Renderable(x) for x in [base] + copies])
    d = np.abs(repeated - three)
    print(f"[render multi] repeater vs copies: alpha mean {three[..., 3].mean():.3f}; |diff| max {d.max():.2e} "
          f"p99.9 {np.percentile(d, 99.9):.2e} mean {d.mean():.2e}", flush=True)
    assert three[..., 3].mean() > 0.02
    # equal but for samples on the shared faces of the copies' boxes, which
    # the copies evaluate twice
    # measured: max 1.8e-5, mean 4e-8
    assert d.mean() < 1e-6 and d.max() < 5e-4
    for x in [base] + copies:
        mgr.destroy(x)


def test_device_render_target_equals_host_target(cuda, tmp_path):
    """CUDARenderBuffer (reference cuda-render-buffer.cuh:11-42): the render
    lands in device memory, equal bit for bit to the host target's image,
    readable in place through its device address, and saved the same way."""
    import torch

    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    mgr = tn.NeRFManager()
    na = _nerf(tn, mgr, scenes.blob_field(61, 8, 1.0), 1.0)
    nb = _nerf(tn, mgr, scenes.blob_field(62, 8, 1.0), 1.0, _rigid(0.3, [0, 1, 0], [0.2, 0, 0]))
    cam = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(63, 1, 2.2, 40, 30, 50.0))[0]
    for renderables in ([tn.Renderable(na)], [tn.Renderable(na), tn.Renderable(nb)]):
        host, dev = tn.CPURenderBuffer(), tn.CUDARenderBuffer()
        r = tn.Renderer()
        r.submit(tn.RenderRequest(cam, renderables, host))
        r.submit(tn.RenderRequest(cam, renderables, dev))
        h = np.array(host.get_rgba())
        assert (dev.width, dev.height) == (40, 30) and h[..., 3].max() > 0.1
        assert np.array_equal(dev.get_rgba(), h)
        # the device image in place, through its address
        view = torch.empty(30 * 40 * 4, dtype=torch.float32, device="cuda")
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        assert hip.hipMemcpy(view.data_ptr(), dev.data_ptr, view.numel() * 4, 3) == 0  # device to device
        assert np.array_equal(view.cpu().numpy().reshape(30, 40, 4), h)
    host.save_image(str(tmp_path / "h.png"))
    dev.save_image(str(tmp_path / "d.png"))
    assert (tmp_path / "h.png").read_bytes() == (tmp_path / "d.png").read_bytes()
    for x in (na, nb):
        mgr.destroy(x)
