"""The render path end to end through Renderer.submit (reference
NeRFRenderingController::submit, src/controllers/nerf-rendering-controller.cu:
50-136, and Renderer::perform_task, src/core/renderer.cu:172-646) on the
large-scene configuration: a box of side 16 with 5 occupancy cascades of
128^3 and 256 seeded Gaussian blobs.

* The blob field itself (a synthetic scene rendered in place of a network)
  through submit, against the oracle's restatement of the reference's render
  loop (fixed steps dt = clamp(t/256, dt_min, 8 dt_min) from the box entry,
  occupancy tested per sample at the cascade its normalised step picks,
  alpha from the normalised step, colour only above alpha 1e-4, stop below
  T = 1e-4). The two loops place samples differently after a skipped empty
  cell (this build marches only through occupied cells); they agree to 2e-3
  per channel (max) and 2e-5 (mean).
* A network in fp16 (MFMA) against the same network in fp32 at 1920x1080, the
  RGBA within 1e-2 max-abs.
* Cancellation from the progress callback: on_cancel runs, on_complete does
  not, and the renderer renders normally afterwards.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AABB, LEVELS, SEED = 16.0, 5, 404
DT_MIN = math.sqrt(3.0) / 1024.0


def large_blobs():
    from turbonerf_amd import scenes

    return scenes.blob_field(SEED, 256, AABB, sigma_range=(0.004, 0.012))


def test_large_scene_blob_field_through_submit_vs_reference_loop(cuda):
    import oracle
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    W, H = 240, 135
    blobs = large_blobs()
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.set_analytic_field(blobs, AABB, 0.01)
    bits = nerf.get_occupancy_bitfield()
    assert 0.2 < float((bits != 0).mean()) < 0.5
    cams = scenes.sphere_cameras(SEED + 1, 2, 24.0, W, H, 50.0)
    r = tn.Renderer()
    n = W * H
    bbox = scenes.bbox(AABB)
    dt = float(np.float32(DT_MIN))
    for ci, cam in enumerate(cams):
        print(f"[render submit] analytic large scene, camera {ci}, {W}x{H}", flush=True)
        buf = tn.CPURenderBuffer()
        done = []
        r.submit(tn.RenderRequest(synthetic.to_pyturbonerf_cameras(tn, [cam])[0], [tn.Renderable(nerf)], buf,
                                  on_complete=lambda: done.append(1)))
        got = np.array(buf.get_rgba()).reshape(n, 4).T
        assert done == [1]
        rr = oracle.render_rays(n, 0, n, cam, bbox)
        # rounding can leave the entry point an ulp outside the box, where the
        # reference loop would find no sample; both loops start inside it
        rr["ori"] = np.clip(rr["ori"], -AABB / 2, AABB / 2).astype(np.float32)
        tmax = np.where(rr["alive"] != 0, rr["tmax"], np.float32(-1)).astype(np.float32)
        ref = oracle.render_fixed_step_analytic(n, n, rr["ori"], rr["dir"], rr["t"], tmax, bbox, bits, LEVELS, 128,
                                                1 / 256, dt, 8 * dt, 1 << 30, blobs, inv_aabb=1.0 / AABB)
        err = np.abs(got - ref)
        print(f"[render submit] camera {ci}: alpha mean {ref[3].mean():.3f} max {ref[3].max():.3f}; |err| max "
              f"{err.max():.2e} p99.9 {np.percentile(err, 99.9):.2e} mean {err.mean():.2e}", flush=True)
        assert ref[3].max() > 0.5 and ref[3].mean() > 0.02
        assert err.max() < 2e-3 and err.mean() < 2e-5
    mgr.destroy(nerf)


def _network_nerf(tn, synthetic, scene_cams, images, precision, params):
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(AABB), scene_cams, images))
    cfg = tn.TrainingConfig()
    cfg.network.precision = precision
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 18)
    tr.load_images()
    tr.train_step()  # a NeRF renders once it has trained a step
    return mgr, nerf, tr


def test_large_scene_fp16_network_render_matches_fp32_at_1080p(cuda):
    import torch

    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import ops, scenes, synthetic

    blobs = large_blobs()
    bits = torch.zeros(128 ** 3, dtype=torch.uint8, device="cuda")
    dens = torch.zeros(LEVELS * 128 ** 3, dtype=torch.float32, device="cuda")
    ops.blob_density_grid(ops.grid_struct(bits, dens, LEVELS), torch.from_numpy(blobs.ravel()).cuda())
    dens_h = dens.cpu().numpy()
    cams = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(SEED, 4, 24.0, 64, 64, 50.0))
    images = np.zeros((4, 64, 64, 4), np.uint8)
    nerfs = [_network_nerf(tn, synthetic, cams, images, p, None) for p in (0, 1)]
    # one set of fp16-representable weights for both: hash entries U(-1, 1),
    # the MLPs as initialised with the density row of W2 scaled by 8 (so the
    # raw density spans several units and parts of the frame are opaque)
    p = nerfs[1][2].get_params()
    rng = np.random.default_rng(SEED)
    n_grid = len(p) - 9408
    p[:n_grid] = rng.uniform(-1.0, 1.0, n_grid)
    p[n_grid + 2048:n_grid + 2048 + 64] *= 8.0
    p = p.astype(np.float16).astype(np.float32)
    for _, nerf, tr in nerfs:
        tr.set_params(p)
        tr.set_occupancy_from_density(dens_h, 0.01)
    W, H = 1920, 1080
    cam = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(SEED + 1, 1, 24.0, W, H, 50.0))[0]
    out = []
    for prec, (_, nerf, tr) in zip((0, 1), nerfs):
        print(f"[render submit] network prec={prec} {W}x{H}", flush=True)
        buf = tn.CPURenderBuffer()
        r = tn.Renderer()
        r.submit(tn.RenderRequest(cam, [tn.Renderable(nerf)], buf))
        out.append(np.array(buf.get_rgba()))
        print(f"[render submit] prec={prec}: {r.last_sample_count} samples, alpha mean {out[-1][..., 3].mean():.3f}",
              flush=True)
    d = np.abs(out[1] - out[0])
    print(f"[render submit] fp16 vs fp32: max {d.max():.2e} mean {d.mean():.2e} p99.99 "
          f"{np.percentile(d, 99.99):.2e}", flush=True)
    assert out[0][..., 3].mean() > 0.02 and out[0][..., 3].max() > 0.5
    assert d.max() < 1e-2
    for _, _, tr in nerfs:
        tr.teardown()


def test_render_cancel_from_progress_callback(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.set_analytic_field(large_blobs(), AABB, 0.01)
    cam = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(SEED + 1, 1, 24.0, 640, 480, 50.0))[0]
    r = tn.Renderer(batch_size=1 << 16)  # tasks of 8192 rays: many progress calls per frame
    events, progress = [], []
    buf = tn.CPURenderBuffer()

    def on_progress(f):
        progress.append(f)
        if len(progress) == 3:
            req.cancel()

    req = tn.RenderRequest(cam, [tn.Renderable(nerf)], buf, on_complete=lambda: events.append("complete"),
                           on_progress=on_progress, on_cancel=lambda: events.append("cancel"))
    r.submit(req)
    assert events == ["cancel"] and len(progress) == 3 and progress[-1] < 1.0
    full = tn.CPURenderBuffer()
    done = []
    r.submit(tn.RenderRequest(cam, [tn.Renderable(nerf)], full, on_complete=lambda: done.append(1)))
    a = np.array(full.get_rgba())
    assert done == [1] and a[..., 3].max() > 0.5
    # the cancelled frame holds the rows rendered before the cancel, zero after
    part = np.array(buf.get_rgba()).reshape(-1, 4)
    m = 3 * 8192
    assert np.array_equal(part[:m], a.reshape(-1, 4)[:m]) and not part[m:].any()
