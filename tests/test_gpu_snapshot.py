"""Snapshot files (`.turbo`, the reference's FileManager format,
src/services/file-manager.cuh:17-198): a NeRF trained for 64 steps is saved;
the file is the 300-byte header (project name, version 0, parameter counts,
bitfield size, aabb), the fp32 parameters and the bitfield, byte for byte what
the trainer holds; loaded into a second NeRF it renders the same image bit for
bit; files that are not snapshots, of another version or truncated are
refused."""
import struct

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _render(tn, cam, nerf):
    buf = tn.CPURenderBuffer()
    tn.Renderer().submit(tn.RenderRequest(cam, [tn.Renderable(nerf)], buf))
    return np.array(buf.get_rgba(), copy=True)


def test_snapshot_round_trip_and_errors(cuda, tmp_path):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import scenes, synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=5, n_cams=8, width=64, height=64)
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
    cfg = tn.TrainingConfig()
    cfg.network.precision = 1
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 16)
    tr.load_images()
    for i in range(4):
        tr.train_steps(16, read_metrics=False)
        tr.update_occupancy_grid(16 * (i + 1))
    tr.synchronize()
    path = str(tmp_path / "scene.turbo")
    tn.FileManager.save(nerf, path)

    raw = open(path, "rb").read()
    name, dpath, version, n_color, n_density, n_app, n_bits, aabb, n_appearances = struct.unpack("<16s256s7I", raw[:300])
    params = np.asarray(tr.get_params(), np.float32)
    bits = np.asarray(tr.get_occupancy_bitfield(), np.uint8)
    assert name.rstrip(b"\0") == b"TurboNeRF" and version == 0 and aabb == 1
    assert n_color == 32 * 64 + 64 * 64 + 3 * 64 and n_app == 0 and n_appearances == 0
    assert n_density + n_color == params.size and n_bits == bits.size == 128 ** 3
    assert len(raw) == 300 + 4 * params.size + bits.size
    assert np.array_equal(np.frombuffer(raw[300:300 + 4 * params.size], np.float32).view(np.uint32),
                          params.view(np.uint32))
    assert np.array_equal(np.frombuffer(raw[300 + 4 * params.size:], np.uint8), bits)

    nerf2 = mgr.create()
    tn.FileManager.load(nerf2, path)
    cam = synthetic.to_pyturbonerf_cameras(tn, scenes.sphere_cameras(9, 1, 2.0, 96, 80, 50.0))[0]
    a, b = _render(tn, cam, nerf), _render(tn, cam, nerf2)
    assert a[..., 3].max() > 0.05
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))

    bad = tmp_path / "bad.turbo"
    bad.write_bytes(b"x" * 400)
    with pytest.raises(RuntimeError, match="not a TurboNeRF file"):
        tn.FileManager.load(mgr.create(), str(bad))
    v1 = tmp_path / "v1.turbo"
    v1.write_bytes(raw[:272] + struct.pack("<I", 1) + raw[276:])
    with pytest.raises(RuntimeError, match="version 1 is not supported"):
        tn.FileManager.load(mgr.create(), str(v1))
    cut = tmp_path / "cut.turbo"
    cut.write_bytes(raw[:len(raw) // 2])
    with pytest.raises(RuntimeError, match="truncated"):
        tn.FileManager.load(mgr.create(), str(cut))
    tr.teardown()
