"""Lens-distorted cameras on the GPU: training rays from 6 sphere cameras
(96x72, 60 degree field of view) under three lens models (barrel, pincushion
with a cubic term, tangential only) bit-exact against the oracle (origins,
directions, t, t_max, targets, image ids, alive flags); render rays of one
camera within 1e-6 of the oracle's (the render kernel renormalises the
direction after the frame change) and away from the pinhole camera's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NC, W, H, R = 6, 96, 72, 8192
LENSES = [(-0.28, 0.07, 0.0, 1e-4, -2e-4),
          (0.12, -0.02, 0.003, -5e-4, 8e-4),
          (0.0, 0.0, 0.0, 2e-3, 0.0)]


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("k", LENSES)
def test_distorted_training_rays_bit_exact(cuda, k):
    import torch

    import oracle
    from turbonerf_amd import ops, scenes

    print(f"[lens train] k={k}", flush=True)
    rng = np.random.default_rng(51)
    cams = scenes.sphere_cameras(52, NC, 2.0, W, H, 60.0, dist_params=k)
    images = rng.integers(0, 256, size=(NC, H, W, 4), dtype=np.uint8)
    rnd = rng.uniform(0, 1, R).astype(np.float32)
    bbox = scenes.bbox(1.0)
    rays = ops.Rays.empty(R)
    ops.init_training_rays(rays, R, NC, W, H, R / NC, ops.upload_struct(bbox),
                           torch.from_numpy(scenes.cameras_bytes(cams)).cuda(), torch.from_numpy(images).cuda(),
                           torch.from_numpy(rnd).cuda())
    rr = oracle.init_training_rays(R, R, NC, W, H, R / NC, bbox, cams, images, rnd)
    alive = rays.alive.cpu().numpy()
    assert np.array_equal(alive, rr["alive"])
    assert alive.sum() > R // 8
    live = alive.astype(bool)
    for key, n in (("ori", 3), ("dir", 3), ("pix_rgba", 4)):
        got = getattr(rays, key).cpu().numpy().reshape(n, R)[:, live]
        want = rr[key].reshape(n, R)[:, live]
        assert np.array_equal(_u32(got), _u32(want)), key
    for key in ("t", "tmax"):
        assert np.array_equal(_u32(getattr(rays, key).cpu().numpy()[live]), _u32(rr[key][live])), key
    assert np.array_equal(rays.img_id.cpu().numpy().view(np.uint32)[live], rr["img_id"][live])


def test_distorted_render_rays(cuda):
    import torch

    import oracle
    from turbonerf_amd import ops, scenes
    from turbonerf_amd._capi import call, ptr, stream_ptr

    n = W * H
    bbox = scenes.bbox(1.0)
    dirs = []
    for k in (LENSES[0], (0.0,) * 5):
        cam = scenes.sphere_cameras(53, 1, 2.0, W, H, 60.0, dist_params=k)[0]
        print(f"[lens render] k={k}", flush=True)
        rays = ops.Rays.empty(n)
        pixel = torch.zeros(n, dtype=torch.int32, device="cuda")
        call("tn_render_rays", n, 0, n, cam, bbox, None, ptr(rays.ori), ptr(rays.dir), ptr(rays.t), ptr(rays.tmax),
             ptr(pixel), ptr(rays.alive), stream_ptr())
        rr = oracle.render_rays(n, 0, n, cam, bbox)
        alive = rays.alive.cpu().numpy().astype(bool)
        assert np.count_nonzero(alive != rr["alive"].astype(bool)) <= 4  # grazing rays may differ
        m = alive & rr["alive"].astype(bool)
        assert m.sum() > n // 8
        d = rays.dir.cpu().numpy().reshape(3, n)
        np.testing.assert_allclose(d[:, m], rr["dir"].reshape(3, n)[:, m], atol=1e-6)
        np.testing.assert_allclose(rays.ori.cpu().numpy().reshape(3, n)[:, m], rr["ori"].reshape(3, n)[:, m],
                                   atol=2e-6)
        dirs.append(d)
    # the barrel lens bends the corner rays by far more than the tolerance
    assert np.abs(dirs[0] - dirs[1]).max() > 1e-2


def test_train_and_render_through_a_lens(cuda):
    """A blob scene photographed through a barrel lens: the trainer and the
    renderer take the PyTurboNeRF cameras with their DistortionParams, and
    after 200 steps the render through the same lens matches the
    photograph as closely as the pinhole scene does (mean |rgb err| < 0.1)."""
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=3, n_cams=16, width=96, height=96, dist_params=LENSES[0])
    pycams = synthetic.to_pyturbonerf_cameras(tn, cams)
    assert abs(pycams[0].dist_params.k1 - LENSES[0][0]) < 1e-7
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), pycams, images))
    tr = tn.Trainer(nerf)
    tr.setup_data(batch_size=1 << 16)
    tr.load_images()
    for i in range(200):
        tr.train_step()
        if i % 16 == 0 and i > 0:
            tr.update_occupancy_grid(i)
    buf = tn.CPURenderBuffer()
    tn.Renderer().submit(tn.RenderRequest(pycams[0], [tn.Renderable(nerf)], buf))
    rgba = np.array(buf.get_rgba())
    gt = images[0].astype(np.float32) / 255.0
    err = np.abs(rgba[..., :3] - gt[..., :3] * gt[..., 3:4]).mean()
    print(f"[lens e2e] mean |rgb err| = {err:.4f}", flush=True)
    assert np.isfinite(rgba).all() and rgba[..., 3].max() > 0.3
    assert err < 0.1
