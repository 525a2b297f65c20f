"""Preview renders of the grid patterns (reference
src/controllers/nerf-rendering-controller.cu:84-131,
src/rendering/ray-batch-coordinators/rectangular-grid-ray-batch-coordinator.cu:62-105).

* A request with the Final flag renders the same image, bit for bit, under
  every pattern: the grid patterns only add the preview in front.
* A Preview-only request under a grid pattern renders about 2^14 rays spread
  over the image and fills each block of pixels with its ray's colour: the
  image equals, bit for bit, a render of the camera scaled to the preview
  grid, expanded to full size by nearest neighbour (output pixel (x, y)
  takes grid pixel (x gw / W, y gh / H)). The hexagonal pattern takes the
  same rectangular grid.
* Progress: the preview reports 1 / (tasks) first, the final tasks end at 1.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(tn):
    from turbonerf_amd import scenes

    aabb = 16.0
    blobs = scenes.blob_field(405, 256, aabb, sigma_range=(0.004, 0.012))
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.set_analytic_field(blobs, aabb, 0.01)
    return mgr, nerf, scenes.sphere_cameras(406, 1, 24.0, 320, 180, 50.0)[0]


def _render(tn, cam, nerf, pattern, flags, progress=None):
    buf = tn.CPURenderBuffer()
    r = tn.Renderer(pattern)
    r.submit(tn.RenderRequest(cam, [tn.Renderable(nerf)], buf, flags=flags, on_progress=progress))
    return np.array(buf.get_rgba())


def _grid(W, H):
    gh = max(1, min(H, int(round(np.sqrt(16384.0 / (W / H))))))
    return max(1, min(W, 16384 // gh)), gh


def test_grid_patterns_final_and_preview(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    mgr, nerf, c = _scene(tn)
    cam = synthetic.to_pyturbonerf_cameras(tn, [c])[0]
    W, H = cam.resolution
    P = tn.RenderPattern
    F = tn.RenderFlags
    ref = _render(tn, cam, nerf, P.LinearChunks, F.Final)
    assert ref[..., 3].mean() > 0.02
    for pat in (P.RectangularGrid, P.HexagonalGrid):
        assert np.array_equal(_render(tn, cam, nerf, pat, F.Final), ref)
        prog = []
        assert np.array_equal(_render(tn, cam, nerf, pat, F.Preview | F.Final, prog.append), ref)
        assert prog[-1] == 1.0 and all(a < b for a, b in zip(prog, prog[1:]))

    gw, gh = _grid(W, H)
    assert (gw, gh) == (170, 96)
    f32 = np.float32
    fx, fy = cam.focal_length
    px, py = cam.principal_point
    small = tn.Camera(resolution=(gw, gh), near=cam.near, far=cam.far,
                      focal_length=(float(f32(fx) * f32(gw) / f32(W)), float(f32(fy) * f32(gh) / f32(H))),
                      principal_point=(float(f32(px) * f32(gw) / f32(W)), float(f32(py) * f32(gh) / f32(H))),
                      shift=cam.shift, transform=cam.transform, dist_params=cam.dist_params)
    low = _render(tn, small, nerf, P.LinearChunks, F.Final).reshape(gh, gw, 4)
    ys, xs = np.arange(H) * gh // H, np.arange(W) * gw // W
    expanded = low[ys[:, None], xs[None, :]]
    for pat in (P.RectangularGrid, P.HexagonalGrid):
        prog = []
        got = _render(tn, cam, nerf, pat, F.Preview, prog.append).reshape(H, W, 4)
        print(f"[preview] {pat}: grid {gw}x{gh}, progress {prog}, |preview - final| mean "
              f"{np.abs(got - ref.reshape(H, W, 4)).mean():.3e}", flush=True)
        assert np.array_equal(got, expanded)
        n_tasks = -(-(W * H) // ((1 << 20) // 8)) + 1
        assert prog == [np.float32(1.0 / n_tasks)]
    # under LinearChunks a Preview-only request has nothing to render
    assert not _render(tn, cam, nerf, P.LinearChunks, F.Preview).any()
    mgr.destroy(nerf)
