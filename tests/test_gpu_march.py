"""GPU parity of the plain-render path (4096 rays from 8 sphere cameras over an
8-blob field, single-cascade 128^3 bitfield thresholded at 0.01, dt = sqrt(3)/1024,
at most 1024 samples per ray): uniforms, rays, bitfield, per-ray counts,
offsets and every per-sample output bit-exact against the oracle; composited
RGB within 2e-3 of the oracle's fp64 compositing."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 1234
N_RAYS = 4096
N_CAMS = 8
RES = 256
DT_MIN = math.sqrt(3.0) / 1024.0
CONE = 1.0 / 256.0
S_CAP = 1 << 20


@pytest.fixture(scope="module")
def scene(cuda):
    import torch

    from turbonerf_amd import ops, scenes

    blobs = scenes.blob_field(SEED, 8, 1.0)
    cams = scenes.sphere_cameras(SEED, N_CAMS, 2.0, RES, RES, 50.0)
    rng = np.random.default_rng(SEED)
    images = rng.integers(0, 256, size=(N_CAMS, RES, RES, 4), dtype=np.uint8)
    bbox = scenes.bbox(1.0)
    dev = dict(
        blobs=torch.from_numpy(blobs.ravel()).cuda(),
        cams=torch.from_numpy(scenes.cameras_bytes(cams)).cuda(),
        bbox=ops.upload_struct(bbox),
        images=torch.from_numpy(images).cuda(),
    )
    return dict(blobs=blobs, cams=cams, images=images, bbox=bbox, dev=dev)


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def test_c0_pipeline_parity(scene):
    import torch

    import oracle
    from turbonerf_amd import ops

    dev = scene["dev"]
    print(f"[c0] seed={SEED} rays={N_RAYS} cams={N_CAMS}", flush=True)
    # uniforms
    rnd = ops.fill_uniform(4 * N_RAYS, SEED)
    rnd_ref = oracle.uniform(4 * N_RAYS, SEED)
    assert np.array_equal(rnd.cpu().numpy().view(np.uint32), rnd_ref.view(np.uint32))

    # bitfield from the analytic field
    vol = 128 ** 3
    bits = torch.zeros(vol, dtype=torch.uint8, device="cuda")
    dens = torch.zeros(vol, dtype=torch.float32, device="cuda")
    grid = ops.grid_struct(bits, dens, 1)
    ops.blob_density_grid(grid, dev["blobs"])
    ops.occ_update_bits(grid, 0.01)
    dens_h = dens.cpu().numpy()
    bits_h = bits.cpu().numpy()
    assert np.array_equal(bits_h, oracle.occ_update_bits(dens_h, vol, 1, 0.01))
    assert 0 < bits_h.sum() < vol
    # cell-centre densities against the oracle's fp64 evaluation
    idx = np.arange(0, vol, 97, dtype=np.uint32)
    xyz = np.stack([[oracle.morton3d_invert(int(m) >> k) for k in range(3)] for m in idx]).astype(np.float32)
    xyz = (xyz - 64.0 + 0.5) / 128.0
    s_ref, _ = oracle.blob_eval(xyz, scene["blobs"])
    np.testing.assert_allclose(dens_h[idx], s_ref, rtol=1e-4, atol=1e-5)

    # rays
    rays = ops.Rays.empty(N_RAYS)
    ops.init_training_rays(rays, N_RAYS, N_CAMS, RES, RES, N_RAYS / N_CAMS, dev["bbox"], dev["cams"],
                           dev["images"], rnd[:N_RAYS])
    rr = oracle.init_training_rays(N_RAYS, N_RAYS, N_CAMS, RES, RES, N_RAYS / N_CAMS, scene["bbox"], scene["cams"],
                                   scene["images"], rnd_ref[:N_RAYS])
    alive = rays.alive.cpu().numpy()
    assert np.array_equal(alive, rr["alive"])
    assert alive.sum() > N_RAYS // 8
    for k in ("ori", "dir", "pix_rgba"):
        assert np.array_equal(getattr(rays, k).cpu().numpy().view(np.uint32), rr[k].view(np.uint32)), k
    for k in ("t", "tmax"):
        assert np.array_equal(getattr(rays, k).cpu().numpy().view(np.uint32), rr[k].view(np.uint32)), k
    assert np.array_equal(_u32(rays.img_id), rr["img_id"])

    # march count (cap 1024 samples per ray)
    ops.march_count(rays, N_RAYS, scene["bbox"], grid, CONE, DT_MIN, DT_MIN, 1024)
    n_ref, alive_ref = oracle.march_count(N_RAYS, N_RAYS, scene["bbox"], bits_h, 1, 128, CONE, DT_MIN, DT_MIN, rr, 1024)
    n_gpu = _u32(rays.n_steps)
    assert np.array_equal(n_gpu, n_ref)
    assert np.array_equal(rays.alive.cpu().numpy(), alive_ref)
    assert n_gpu.max() <= 1024 and n_gpu.sum() > 0
    print(f"[c0] samples={int(n_gpu.sum())} max/ray={int(n_gpu.max())}", flush=True)

    # scan: training selection (mode 0) against the oracle
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    budget = int(n_gpu.sum()) // 2
    ops.scan_select(rays, N_RAYS, budget, info, mode=0)
    offs_ref, nr_ref, ns_ref, rc = oracle.scan_select(n_ref, N_RAYS, budget)
    inf = _u32(info)
    assert np.array_equal(_u32(rays.offset), offs_ref)
    assert (inf[0], inf[1], inf[2]) == (nr_ref, ns_ref, rc)
    # render selection (mode 1): every ray fits
    ops.scan_select(rays, N_RAYS, S_CAP, info, mode=1)
    inf = _u32(info)
    excl = np.concatenate([[0], np.cumsum(n_ref)[:-1]]).astype(np.uint32)
    assert np.array_equal(_u32(rays.offset), excl)
    assert inf[0] == N_RAYS and inf[1] == n_ref.sum() and inf[2] == 0

    # generate
    samples = ops.Samples.empty(S_CAP)
    ops.march_generate(rays, samples, N_RAYS, info, scene["bbox"], 1.0, grid, DT_MIN, DT_MIN, CONE)
    so = oracle.march_generate(N_RAYS, N_RAYS, S_CAP, scene["bbox"], 1.0, bits_h, 1, 128, DT_MIN, DT_MIN, CONE, rr, excl,
                               alive_ref, n_ref)
    ns = int(n_ref.sum())
    for k in ("t", "dt", "m_norm", "dt_norm"):
        assert np.array_equal(getattr(samples, k)[:ns].cpu().numpy().view(np.uint32), so[k][:ns].view(np.uint32)), k
    for k in ("pos", "dir"):
        g = getattr(samples, k).cpu().numpy().reshape(3, S_CAP)[:, :ns]
        assert np.array_equal(g.view(np.uint32), so[k].reshape(3, S_CAP)[:, :ns].view(np.uint32)), k
    assert np.array_equal(_u32(samples.img_id)[:ns], so["img_id"][:ns])

    # analytic compositing
    sigma, rgb = ops.blob_eval(info, S_CAP, S_CAP, 1.0, dev["blobs"], samples.pos)
    rgba = ops.composite_forward(info, S_CAP, S_CAP, N_RAYS, rays.n_steps, rays.offset, sigma, rgb, samples.dt,
                                 raw_density=False).cpu().numpy().reshape(4, N_RAYS)
    ref = oracle.composite_analytic(N_RAYS, n_ref, excl, so["pos"], S_CAP, 1.0, so["dt"], scene["blobs"])
    err = np.abs(rgba - ref).max()
    print(f"[c0] composite max|err| vs fp64 = {err:.3e}", flush=True)
    assert err < 2e-3
    assert ref[3].max() > 0.5  # some rays are nearly opaque
