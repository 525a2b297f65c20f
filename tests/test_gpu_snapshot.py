"""Snapshot files (`.turbo`, the reference's FileManager format,
src/services/file-manager.cuh:17-198): a NeRF trained for 64 steps is saved;
the file is the 300-byte header (project name, version 0, parameter counts,
bitfield size, aabb), the fp32 parameters (the colour network's output layer
padded to 16 zero rows, tiny-cuda-nn's count) and the bitfield, byte for byte
what the trainer holds; loaded into a second NeRF it renders the same image bit
for bit; files that are not snapshots, of another version, truncated or with a
bitfield of another size are refused, the last before the NeRF changes.
With the optimiser state (reference NGPAdamOptimizer::serialize,
src/core/adam-optimizer.cuh:203-224) a load restores the moments, the
per-parameter step counts and the training step exactly, and the next step
continues from them."""
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
    assert n_color == 32 * 64 + 64 * 64 + 16 * 64 and n_app == 0 and n_appearances == 0
    assert n_density + 32 * 64 + 64 * 64 + 3 * 64 == params.size and n_bits == bits.size == 128 ** 3
    n_stored = n_density + n_color
    assert len(raw) == 300 + 4 * n_stored + bits.size
    stored = np.frombuffer(raw[300:300 + 4 * n_stored], np.float32)
    assert np.array_equal(stored[:params.size].view(np.uint32), params.view(np.uint32))
    assert not stored[params.size:].any()  # the 13 zero rows of the padded output layer
    assert np.array_equal(np.frombuffer(raw[300 + 4 * n_stored:], np.uint8), bits)

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
    # a bitfield of another size is refused before the NeRF changes
    small = tmp_path / "small.turbo"
    small.write_bytes(raw[:288] + struct.pack("<I", 64 ** 3) + raw[292:300 + 4 * n_stored] + bits[:64 ** 3].tobytes())
    with pytest.raises(RuntimeError, match="bitfield"):
        tn.FileManager.load(nerf2, str(small))
    assert np.array_equal(_render(tn, cam, nerf2).view(np.uint32), b.view(np.uint32))
    tr.teardown()


def test_snapshot_with_optimizer_state_resumes_exactly(cuda, tmp_path):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=6, n_cams=8, width=64, height=64)
    tcams = synthetic.to_pyturbonerf_cameras(tn, cams)

    def trainer():
        mgr = tn.NeRFManager()
        nerf = mgr.create()
        nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), tcams, images))
        cfg = tn.TrainingConfig()
        cfg.network.precision = 1
        cfg.network.beta2 = 0.99
        tr = tn.Trainer(nerf, cfg)
        tr.setup_data(batch_size=1 << 16)
        tr.load_images()
        return mgr, nerf, tr

    mgr, nerf, tr = trainer()
    for i in range(3):
        tr.train_steps(16, read_metrics=False)
        tr.update_occupancy_grid(16 * (i + 1))
    tr.synchronize()
    path = str(tmp_path / "resume.turbo")
    tn.FileManager.save(nerf, path, include_optimizer_state=True)
    saved = tr._get_optimizer_state()
    params = tr.get_params()
    bits = tr.get_occupancy_bitfield()
    raw = open(path, "rb").read()
    n = params.size
    n_plain = 300 + 4 * (n + 13 * 64) + 128 ** 3
    assert raw[n_plain:n_plain + 8] == b"TNADAM01" and len(raw) == n_plain + 16 + 12 * n
    assert struct.unpack("<2I", raw[n_plain + 8:n_plain + 16]) == (48, n)

    mgr2, nerf2, tr2 = trainer()
    # a file without the section loads with a fresh optimiser state
    plain = tmp_path / "plain.turbo"
    plain.write_bytes(raw[:n_plain])
    assert tn.FileManager.load(nerf2, str(plain)) is False
    assert not tr2._get_optimizer_state()["m1"].any() and nerf2.training_step == 0
    assert tn.FileManager.load(nerf2, path) is True
    st = tr2._get_optimizer_state()
    assert nerf2.training_step == nerf.training_step == 48
    for k in ("m1", "m2", "steps"):
        assert np.array_equal(st[k].view(np.uint32), saved[k].view(np.uint32)), k
    assert np.array_equal(tr2.get_params().view(np.uint32), params.view(np.uint32))
    assert np.array_equal(tr2.get_occupancy_bitfield(), bits)
    # the next step continues from the loaded moments: for every parameter
    # the step touched, m1' = b1 m1 + (1 - b1) g and m2' = b2 m2 + (1 - b2) g^2
    # with one g, and its step count went up by one
    tr2.train_step()
    after = tr2._get_optimizer_state()
    b1, b2 = np.float32(0.9), np.float32(0.99)
    touched = after["steps"] != saved["steps"]
    assert touched.sum() > 9408 and np.all(after["steps"][touched] == saved["steps"][touched] + 1)
    m1s, m2s = saved["m1"][touched].astype(np.float64), saved["m2"][touched].astype(np.float64)
    g = (after["m1"][touched].astype(np.float64) - float(b1) * m1s) / (1 - float(b1))
    g2 = (after["m2"][touched].astype(np.float64) - float(b2) * m2s) / (1 - float(b2))
    # entries where the step's share is not buried under the old moments'
    # rounding, and where the old moments were not zero (a reset would pass)
    clear = ((1 - float(b1)) * np.abs(g) > 0.1 * np.abs(m1s)) & ((1 - float(b2)) * g * g > 0.1 * m2s) & (m2s > 0)
    assert clear.sum() > 1000
    np.testing.assert_allclose(g2[clear], g[clear] ** 2, rtol=1e-3)
    assert nerf2.training_step == 49
    tr.teardown()
    tr2.teardown()
