"""Training from files on the GPU: the seeded blob scene (16 cameras, 96x96)
written as PNG images and a transforms.json (NeRF-convention matrices,
focal lengths and principal points per frame), loaded with
Dataset(file_path).load_transforms() and Trainer.load_images(). The first
training step gives the same loss, bit for bit, as the same scene handed over
in memory (same cameras and pixels on the device); training then converges as
the in-memory scene does (loss halves in 120 steps)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]], np.float32)
F = np.diag([1, -1, -1, 1]).astype(np.float32)


def _write_scene(tmp_path, cams, images):
    from PIL import Image

    frames = []
    for i, c in enumerate(cams):
        Image.fromarray(images[i], "RGBA").save(tmp_path / f"img_{i:03d}.png")
        m = np.array(c.transform[:], np.float32).reshape(3, 4)
        nerf = np.eye(4, dtype=np.float32)
        nerf[:3] = P.T @ m @ F
        frames.append({"file_path": f"img_{i:03d}.png", "transform_matrix": nerf.astype(np.float64).tolist(),
                       "fl_x": float(c.focal_length[0]), "fl_y": float(c.focal_length[1]),
                       "cx": float(c.principal_point[0]), "cy": float(c.principal_point[1])})
    doc = {"w": int(cams[0].resolution[0]), "h": int(cams[0].resolution[1]), "aabb_scale": 1,
           "near": float(cams[0].near_plane), "far": float(cams[0].far_plane), "frames": frames}
    path = tmp_path / "transforms.json"
    path.write_text(json.dumps(doc))
    return str(path)


def _trainer(tn, ds):
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(ds)
    cfg = tn.TrainingConfig()
    cfg.network.precision = 1
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 16)
    tr.load_images()
    return mgr, nerf, tr


def test_train_from_transforms_json(cuda, tmp_path):
    pytest.importorskip("PIL")
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=3, n_cams=16, width=96, height=96)
    path = _write_scene(tmp_path, cams, images)
    ds_file = tn.Dataset(path)
    ds_file.load_transforms()
    assert ds_file.n_images == 16 and ds_file.image_dimensions == (96, 96)
    ds_mem = tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images)
    for a, b in zip(ds_file.cameras, ds_mem.cameras):
        np.testing.assert_array_equal(a.transform.to_matrix(), b.transform.to_matrix())
        assert (a.focal_length, a.principal_point, a.near, a.far) == (b.focal_length, b.principal_point, b.near,
                                                                      b.far)
    print(f"[files] {path}", flush=True)
    mgr_mem, nerf_mem, tr_mem = _trainer(tn, ds_mem)
    mgr, nerf, tr_file = _trainer(tn, ds_file)
    m_mem, m_file = tr_mem.train_step(), tr_file.train_step()
    print(f"[files] first step loss file {m_file.loss!r} memory {m_mem.loss!r}", flush=True)
    assert m_file.loss == m_mem.loss and m_file.n_rays == m_mem.n_rays and m_file.n_samples == m_mem.n_samples
    losses = [m_file.loss]
    for i in range(1, 120):
        losses.append(tr_file.train_step().loss)
        if i % 16 == 0:
            tr_file.update_occupancy_grid(i)
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    print(f"[files] loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.5 * first and nerf.can_render
