"""Large scene with cascades: AABB side 16 (5 cascades of 128^3), a field of
256 small Gaussian blobs. On the GPU against the oracle:

* occupancy update: decay by 0.95 and the jittered sample points of every
  cell of a cascade (2^21 points) bit-exact; the max-update with both the
  warm-up rule (all cells) and the random-subset rule within a few ulp (the
  density activation goes through the device exp); threshold 0.01, the
  5-bit bitfield and its popcount bit-exact given identical densities;
* training rays through the cascades (dt from sqrt(3)/1024 to 16x that,
  cone 1/256): per-ray counts, offsets and every per-sample output bit-exact,
  with few rays (a wave per ray) and with many (lanes skip empty space for
  their own rays).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AABB = 16.0
LEVELS = 5
RES = 128
VOL = RES ** 3
DT_MIN = math.sqrt(3.0) / 1024.0
CONE = 1.0 / 256.0
SEED = 404


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.fixture(scope="module")
def field(cuda):
    import torch

    from turbonerf_amd import ops, scenes

    # small blobs (0.06-0.2 world units) so every cascade is partly empty
    blobs = scenes.blob_field(SEED, 256, AABB, sigma_range=(0.004, 0.012))
    bits = torch.zeros(VOL, dtype=torch.uint8, device="cuda")
    dens = torch.zeros(VOL * LEVELS, dtype=torch.float32, device="cuda")
    grid = ops.grid_struct(bits, dens, LEVELS, RES)
    ops.blob_density_grid(grid, torch.from_numpy(blobs.ravel()).cuda())
    ops.occ_update_bits(grid, 0.01)
    return dict(blobs=blobs, bits=bits, dens=dens, grid=grid)


def test_cascade_bitfield_and_update(field):
    import torch

    import oracle
    from turbonerf_amd import ops

    dens0 = field["dens"].cpu().numpy()
    bits0 = field["bits"].cpu().numpy()
    assert np.array_equal(bits0, oracle.occ_update_bits(dens0, VOL, LEVELS, 0.01))
    occ_levels = [int(((bits0 >> k) & 1).sum()) for k in range(LEVELS)]
    print(f"[cascades] occupied cells per cascade {occ_levels}", flush=True)
    assert all(0 < c < VOL for c in occ_levels[1:])  # cascade 0 (the central unit box) may be empty
    assert ops.popcount(field["bits"]) == oracle.popcount(bits0)

    # one update: decay, then per cascade jittered points and a max-update
    # from synthetic raw densities (the network's output), then the bits
    bits = field["bits"].clone()
    dens = field["dens"].clone()
    grid = ops.grid_struct(bits, dens, LEVELS, RES)
    ops.occ_decay(grid, 0.95)
    d_ref = oracle.occ_decay(dens0, LEVELS, 0.95)
    assert np.array_equal(dens.cpu().numpy(), d_ref)
    rng = np.random.default_rng(SEED)
    bits_ref = bits0.copy()
    for level in range(LEVELS):
        rnd = ops.fill_uniform(4 * VOL, SEED + level)
        rnd_h = oracle.uniform(4 * VOL, SEED + level)
        pos = torch.zeros(3 * VOL, dtype=torch.float32, device="cuda")
        ops.occ_sample_points(VOL, VOL, 0, grid, level, 1.0 / AABB, rnd, pos)
        pos_ref = oracle.occ_sample_points(VOL, VOL, 0, RES, level, 1.0 / AABB, rnd_h)
        assert np.array_equal(pos.cpu().numpy().view(np.uint32), pos_ref.view(np.uint32)), level
        raw = rng.normal(-2.0, 4.0, VOL).astype(np.float32)
        sample_all = level % 2 == 0  # the warm-up rule on even cascades, the random subset on odd ones
        ops.occ_update_density(VOL, 0, grid, level, sample_all, rnd[3 * VOL:], torch.from_numpy(raw).cuda())
        d_ref = oracle.occ_update_density(VOL, 0, VOL, level, sample_all, rnd_h[3 * VOL:], raw, bits_ref, d_ref)
    # sigma = exp(min(s, 12) - 1) goes through the device's exp, so the
    # updated densities agree to a few ulp, and the cells each rule touched
    # agree exactly (up to ties decided by those ulps)
    d_gpu = dens.cpu().numpy()
    np.testing.assert_allclose(d_gpu, d_ref, rtol=4e-6, atol=0)
    d_decayed = oracle.occ_decay(dens0, LEVELS, 0.95)
    assert np.count_nonzero((d_gpu != d_decayed) != (d_ref != d_decayed)) <= 16
    # the bitfield is bit-exact given identical densities
    ops.occ_update_bits(grid, 0.01)
    bits_ref = oracle.occ_update_bits(d_gpu, VOL, LEVELS, 0.01)
    assert np.array_equal(bits.cpu().numpy(), bits_ref)
    assert ops.popcount(bits) == oracle.popcount(bits_ref)


@pytest.mark.parametrize("n_rays,max_steps", [(2048, 0xFFFFFFFF), (40960, 0xFFFFFFFF), (40960, 48)])
def test_cascade_marching(field, n_rays, max_steps):
    """2048 rays: a wave marches one ray at a time. 40960 rays (five per wave
    of the bounded grid): lanes take their own rays through empty cells and
    the wave takes each ray's occupied runs together; also with a cap of 48
    samples per ray, which ends runs part-way. All bit-exact."""
    import torch

    import oracle
    from turbonerf_amd import ops, scenes

    n_cams, W = 8, 128
    cams = scenes.sphere_cameras(SEED, n_cams, 0.9 * AABB, W, W, 60.0)
    rng = np.random.default_rng(SEED)
    images = rng.integers(0, 256, size=(n_cams, W, W, 4), dtype=np.uint8)
    bbox = scenes.bbox(AABB)
    rnd = ops.fill_uniform(n_rays, SEED)
    rays = ops.Rays.empty(n_rays)
    ops.init_training_rays(rays, n_rays, n_cams, W, W, n_rays / n_cams, ops.upload_struct(bbox),
                           torch.from_numpy(scenes.cameras_bytes(cams)).cuda(), torch.from_numpy(images).cuda(), rnd)
    rr = oracle.init_training_rays(n_rays, n_rays, n_cams, W, W, n_rays / n_cams, bbox, cams, images,
                                   oracle.uniform(n_rays, SEED))
    assert np.array_equal(rays.alive.cpu().numpy(), rr["alive"])
    bits_h = field["bits"].cpu().numpy()
    dt_max = AABB * DT_MIN
    ops.march_count(rays, n_rays, bbox, field["grid"], CONE, DT_MIN, dt_max, max_steps)
    n_ref, alive_ref = oracle.march_count(n_rays, n_rays, bbox, bits_h, LEVELS, RES, CONE, DT_MIN, dt_max, rr,
                                          max_steps)
    n_gpu = _u32(rays.n_steps)
    assert np.array_equal(n_gpu, n_ref)
    assert np.array_equal(rays.alive.cpu().numpy(), alive_ref)
    ns = int(n_ref.sum())
    print(f"[cascades] rays alive {int(alive_ref.sum())} samples {ns} max/ray {int(n_ref.max())}", flush=True)
    assert ns > 0
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    cap = ns + 64
    ops.scan_select(rays, n_rays, cap, info, mode=1)
    excl = np.concatenate([[0], np.cumsum(n_ref)[:-1]]).astype(np.uint32)
    assert np.array_equal(_u32(rays.offset), excl) and _u32(info)[1] == ns
    smp = ops.Samples.empty(cap)
    ops.march_generate(rays, smp, n_rays, info, bbox, 1.0 / AABB, field["grid"], DT_MIN, dt_max, CONE)
    so = oracle.march_generate(n_rays, n_rays, cap, bbox, 1.0 / AABB, bits_h, LEVELS, RES, DT_MIN, dt_max, CONE, rr,
                               excl, alive_ref, n_ref)
    for k in ("t", "dt", "m_norm", "dt_norm"):
        assert np.array_equal(getattr(smp, k)[:ns].cpu().numpy().view(np.uint32), so[k][:ns].view(np.uint32)), k
    for k in ("pos", "dir"):
        g = getattr(smp, k).cpu().numpy().reshape(3, cap)[:, :ns]
        assert np.array_equal(g.view(np.uint32), so[k].reshape(3, cap)[:, :ns].view(np.uint32)), k
    # the samples visit more than one cascade (dt grows with t)
    dts = np.unique(so["dt"][:ns])
    assert len(dts) > 2
