"""Pins the oracle (oracle/tn_oracle.cpp) to facts independent of its own
formulas: Random123 known-answer vectors, textbook bit interleaving, exact
rational arithmetic, brute-force marching with numpy float32, closed forms of
volume rendering, FP64 finite differences, the O(N^2) mip-NeRF 360 distortion
sum, the Adam closed form, partition of unity / adjointness of the hash grid,
SH orthonormality by exact quadrature, the camera model run backwards, the
lens model against its documented forward form and finite differences, Morton
de-interleaving of occupancy points and known answers at the occupancy update
rule's boundaries. The reference ships no tests or golden
vectors for this path, so these checks are what pin each oracle function."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

F32 = np.float32


# --------------------------------------------------------------- Philox
def _philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    oracle.lib().oc_philox4x32_10(c, k, o)
    return list(o)


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert _philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_stream_definition():
    seed = 0x1234567890AB
    u = oracle.uniform(37, seed, offset=5)
    for j in (0, 1, 2, 3, 17, 36):
        g = 5 + j
        words = _philox([(g >> 2) & 0xFFFFFFFF, g >> 34, 0, 0], [seed & 0xFFFFFFFF, seed >> 32])
        assert u[j] == F32((words[g & 3] >> 8) * 2.0 ** -24)
    assert u.min() >= 0 and u.max() < 1


# --------------------------------------------------------------- Morton
def test_morton_textbook():
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(200, 3)):
        m = 0
        for b in range(10):
            m |= ((int(x) >> b) & 1) << (3 * b) | ((int(y) >> b) & 1) << (3 * b + 1) | ((int(z) >> b) & 1) << (3 * b + 2)
        assert oracle.morton3d(int(x), int(y), int(z)) == m
        assert oracle.morton3d_invert(m) == x and oracle.morton3d_invert(m >> 1) == y and oracle.morton3d_invert(m >> 2) == z


# --------------------------------------------------------------- directed division
def test_fdiv_ru_is_upward_rounding():
    rng = np.random.default_rng(1)
    dt = F32(math.sqrt(3) / 1024)
    vals = np.concatenate([rng.uniform(1e-3, 0.2, 300).astype(F32), [dt, 2 * dt, F32(3) * dt, F32(0.0625)]])
    for a in vals:
        for b in (dt, F32(0.0078125), F32(0.3)):
            q = F32(oracle.fdiv_ru(float(a), float(b)))
            exact = Fraction(float(a)) / Fraction(float(b))
            assert Fraction(float(q)) >= exact
            below = np.nextafter(q, F32(-np.inf))
            assert Fraction(float(below)) < exact


# --------------------------------------------------------------- occupancy queries
def test_grid_level_and_index_bruteforce():
    rng = np.random.default_rng(2)
    for _ in range(300):
        n_levels = int(rng.integers(1, 6))
        p = rng.uniform(-2 ** (n_levels - 1), 2 ** (n_levels - 1), 3).astype(F32)
        dt = F32(rng.uniform(0.001, 0.05))
        lvl = None
        for k in range(n_levels):
            half = F32(0.5) * F32(2 ** k)
            if all(abs(c) < half for c in p) and F32(2 ** k) * F32(1 / 128) > dt:
                lvl = k
                break
        lvl = n_levels - 1 if lvl is None else lvl
        assert oracle.lib().oc_grid_level_at(n_levels, 128, *map(float, p), float(dt)) == lvl
        s = F32(1.0) / F32(2 ** lvl)
        idx = [min(int(max((c * s + F32(0.5)) * F32(128), F32(0))), 127) for c in p]
        assert oracle.lib().oc_voxel_morton_index(128, lvl, *map(float, p)) == oracle.morton3d(*idx)


def _rays(n, seed):
    from turbonerf_amd import scenes

    cams = scenes.sphere_cameras(seed, 4, 2.0, 64, 64, 50.0)
    images = np.zeros((4, 64, 64, 4), np.uint8)
    rnd = oracle.uniform(n, seed)
    return oracle.init_training_rays(n, n, 4, 64, 64, n / 4, scenes.bbox(1.0), cams, images, rnd)


def test_deactivate_rays_known_answers():
    """The reference's alpha-threshold rule (training-batch-kernels.cuh:
    189-208): a ray is dropped iff alpha < threshold and its uniform >
    probability; equality keeps it; rays past n_rays are untouched."""
    B = 8
    alpha = np.array([0.1, 0.1, 0.5, 0.5, 0.49, 0.9, 0.0, 0.1], np.float32)
    rnd = np.array([0.9, 0.5, 0.9, 0.1, 0.51, 0.99, 0.5000001, 0.9], np.float32)
    pix = np.zeros(4 * B, np.float32)
    pix[3 * B:] = alpha
    alive = np.ones(B, np.uint8)
    alive[5] = 0
    got = oracle.deactivate_rays(7, B, 0.5, 0.5, pix, rnd, alive)
    assert got.tolist() == [0, 1, 1, 1, 0, 0, 0, 1]


def test_march_count_all_occupied_bruteforce():
    """All cells occupied, aabb = 1 (constant dt): the count is the number of
    float32 steps t += dt taken while t < tmax and the point stays in the box."""
    from turbonerf_amd import scenes

    n = 256
    rr = _rays(n, 3)
    bits = np.full(128 ** 3, 0xFF, np.uint8)
    dt = F32(math.sqrt(3) / 1024)
    steps, alive = oracle.march_count(n, n, scenes.bbox(1.0), bits, 1, 128, 1 / 256, float(dt), float(dt), rr)
    for i in range(n):
        if not rr["alive"][i]:
            assert steps[i] == 0
            continue
        o = rr["ori"].reshape(3, n)[:, i]
        d = rr["dir"].reshape(3, n)[:, i]
        t, k = rr["t"][i], 0
        while t < rr["tmax"][i]:
            x = o + t * d
            if np.any(x < F32(-0.5)) or np.any(x > F32(0.5)):
                break
            t = F32(t + dt)
            k += 1
        assert steps[i] == k
    bits0 = np.zeros(128 ** 3, np.uint8)
    s0, a0 = oracle.march_count(n, n, scenes.bbox(1.0), bits0, 1, 128, 1 / 256, float(dt), float(dt), rr)
    assert s0.sum() == 0 and a0.sum() == 0


def test_scan_select_rule():
    rng = np.random.default_rng(4)
    steps = rng.integers(0, 50, 1000).astype(np.uint32)
    budget = 9000
    offs, nr, ns, rc = oracle.scan_select(steps, 1000, budget)
    excl = np.concatenate([[0], np.cumsum(steps)[:-1]])
    assert np.array_equal(offs, excl)
    j = np.nonzero(excl < budget)[0].max()
    assert (nr, ns, rc) == (j, excl[j], 0)
    assert ns < budget


def test_march_generate_invariants():
    from turbonerf_amd import scenes

    n = 512
    rr = _rays(n, 5)
    blobs = scenes.blob_field(5, 8, 1.0)
    ii = np.arange(128 ** 3, dtype=np.uint32)
    inv = lambda m: sum(((m >> (3 * b)) & 1) << b for b in range(10))
    xyz = (np.stack([inv(ii), inv(ii >> 1), inv(ii >> 2)], 1).astype(F32) - 64 + 0.5) / 128
    s, _ = oracle.blob_eval(xyz, blobs)
    bits = oracle.occ_update_bits(s, 128 ** 3, 1, 0.01)
    dt = float(F32(math.sqrt(3) / 1024))
    steps, alive = oracle.march_count(n, n, scenes.bbox(1.0), bits, 1, 128, 1 / 256, dt, dt, rr)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    S = int(steps.sum())
    so = oracle.march_generate(n, n, S, scenes.bbox(1.0), 1.0, bits, 1, 128, dt, dt, 1 / 256, rr, offs, alive, steps)
    pos = so["pos"].reshape(3, S)
    for i in np.nonzero(steps)[0]:
        sl = slice(offs[i], offs[i] + steps[i])
        t = so["t"][sl]
        assert np.all(np.diff(t) > 0)
        m = so["m_norm"][sl]
        assert m.min() >= 0 and m.max() <= 1 and np.all(np.diff(m) >= 0)
        # the span runs from the ray's entry t (before any empty-space skip)
        span = max(float(F32(t[-1] + F32(dt)) - rr["t"][i]), 1e-6)
        np.testing.assert_allclose(so["dt_norm"][sl] * span, dt, rtol=1e-5)
        # every sample lies in an occupied cell
        o = rr["ori"].reshape(3, n)[:, i]
        d = rr["dir"].reshape(3, n)[:, i]
        for tt in t[:: max(1, len(t) // 5)]:
            x = o + tt * d
            assert bits[oracle.lib().oc_voxel_morton_index(128, 0, *map(float, x))] & 1
        np.testing.assert_allclose(pos[:, sl[0] if False else offs[i]], np.clip(o + t[0] * d + 0.5, 0, 1), atol=1e-6)


# --------------------------------------------------------------- render loop
def _render_cam(w=40, h=32):
    from turbonerf_amd import scenes

    return scenes.sphere_cameras(9, 2, 2.0, w, h, 50.0)[0]


def test_render_rays_pinhole_and_slab():
    """Pixel rays against a float64 pinhole model and slab intersection: the
    origin moves to the box entry (at least 1e-4 along the ray), t = 0, and
    t_max = min(exit, far - near) from the camera origin."""
    from turbonerf_amd import scenes

    W, H = 40, 32
    cam = _render_cam(W, H)
    n = W * H
    rr = oracle.render_rays(n, 0, n, cam, scenes.bbox(1.0))
    M = np.array(cam.transform[:], np.float64).reshape(3, 4)
    ix, iy = np.arange(n) % W, np.arange(n) // W
    v = np.stack([(ix + cam.offset_px[0] + 0.5) / cam.focal_length[0],
                  (iy + cam.offset_px[1] + 0.5) / cam.focal_length[1], np.ones(n)])
    d = M[:, :3] @ v
    d /= np.linalg.norm(d, axis=0)
    o = M[:, 3:4] + cam.near_plane * np.linalg.norm(v, axis=0) * d
    t_a, t_b = (-0.5 - o) / d, (0.5 - o) / d
    t0, t1 = np.minimum(t_a, t_b).max(0), np.maximum(t_a, t_b).min(0)
    hit = t1 >= t0
    clear = np.abs(t1 - t0) > 1e-4  # skip grazing rays
    assert np.array_equal(rr["alive"][clear].astype(bool), hit[clear])
    assert 0 < hit.sum() < n
    t0c = np.maximum(t0, 1e-4)
    m = hit & clear
    np.testing.assert_allclose(rr["dir"].reshape(3, n)[:, m], d[:, m], atol=1e-6)
    np.testing.assert_allclose(rr["ori"].reshape(3, n)[:, m], (o + t0c * d)[:, m], atol=1e-5)
    np.testing.assert_allclose(rr["tmax"][m], np.minimum(t1, cam.far_plane - cam.near_plane)[m], atol=1e-5)
    assert np.all(rr["t"] == 0)


def _fixed_steps(rr, i, dt, n):
    """float32 brute force of the fixed-step loop: samples inside the box."""
    o = rr["ori"].reshape(3, n)[:, i]
    d = rr["dir"].reshape(3, n)[:, i]
    t, k = F32(0), 0
    while t <= rr["tmax"][i]:
        x = o + t * d
        if np.all(x >= F32(-0.5)) and np.all(x <= F32(0.5)):
            k += 1
        t = F32(t + F32(dt))
    return k


def test_render_fixed_step_closed_forms():
    """A blob 1000x wider than the box is a constant density p with colour c:
    after k in-box samples of constant dt, A = 1 - exp(-p k dt) and rgb = c A;
    below alpha 1e-4 a sample adds alpha but no colour; an opaque first sample
    ends the ray (T < 1e-4); an empty grid renders nothing."""
    from turbonerf_amd import scenes

    W, H = 8, 8
    cam = _render_cam(W, H)
    n = W * H
    bbox = scenes.bbox(1.0)
    rr = oracle.render_rays(n, 0, n, cam, bbox)
    tmax = np.where(rr["alive"] != 0, rr["tmax"], F32(-1)).astype(F32)
    full = np.full(128 ** 3, 0xFF, np.uint8)
    dt = float(F32(math.sqrt(3) / 1024))
    c = np.array([0.2, 0.5, 0.9])

    def render(p, bits=full):
        blobs = np.array([[0, 0, 0, 1000.0, p, *c]], F32)
        return oracle.render_fixed_step_analytic(n, n, rr["ori"], rr["dir"], rr["t"], tmax, bbox, bits, 1, 128,
                                                 1 / 256, dt, dt, 1 << 16, blobs)

    k = np.array([_fixed_steps(rr, i, dt, n) if rr["alive"][i] else 0 for i in range(n)])
    assert k.max() > 300
    for p in (2.0, 0.01):
        out = render(p)
        A = 1 - np.exp(-float(F32(p)) * k * dt)
        np.testing.assert_allclose(out[3], A, rtol=1e-5, atol=1e-12)
        if p > 1:  # alpha 3.4e-3 per sample: colour counts
            np.testing.assert_allclose(out[:3], c[:, None].astype(F32) * A, rtol=1e-5, atol=1e-12)
        else:  # alpha 1.7e-5 per sample: alpha only
            assert np.all(out[:3] == 0) and out[3].max() > 1e-3
    out = render(1e4)
    a1 = 1 - math.exp(-1e4 * dt)
    hitk = k > 0
    np.testing.assert_allclose(out[3][hitk], a1, rtol=1e-6)
    assert np.all(out[3][~hitk] == 0)
    assert np.all(render(2.0, np.zeros(128 ** 3, np.uint8)) == 0)


def test_render_fixed_step_matches_marched_compositing():
    """With every cell occupied and a constant dt the reference's fixed-step
    loop visits the same samples as the marcher, so the two compositings of
    the blob field agree (up to the float32 rounding of normalised positions)."""
    from turbonerf_amd import scenes

    n = 256
    rr = _rays(n, 21)
    bbox = scenes.bbox(1.0)
    full = np.full(128 ** 3, 0xFF, np.uint8)
    dt = float(F32(math.sqrt(3) / 1024))
    blobs = scenes.blob_field(21, 8, 1.0)
    steps, alive = oracle.march_count(n, n, bbox, full, 1, 128, 1 / 256, dt, dt, rr, 4096)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    S = int(steps.sum())
    so = oracle.march_generate(n, n, S, bbox, 1.0, full, 1, 128, dt, dt, 1 / 256, rr, offs, alive, steps)
    ref = oracle.composite_analytic(n, steps, offs, so["pos"], S, 1.0, so["dt"], blobs)
    tmax = np.where(rr["alive"] != 0, rr["tmax"], F32(-1)).astype(F32)
    out = oracle.render_fixed_step_analytic(n, n, rr["ori"], rr["dir"], rr["t"], tmax, bbox, full, 1, 128, 1 / 256,
                                            dt, dt, 4096, blobs)
    assert ref[3].max() > 0.5
    np.testing.assert_allclose(out, ref, rtol=0, atol=2e-6)


# --------------------------------------------------------------- compositing
def _comp(steps, s, c, dt, bg, target, min_T=1e-4):
    n_rays = len(steps)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    ns = int(np.sum(steps))
    return oracle.composite_train(n_rays, n_rays, ns, np.asarray(steps, np.uint32), offs, bg, target, s, c, dt,
                                  1.0, min_T, True)


def test_composite_closed_forms():
    # one sample, alpha = 1 - exp(-sigma dt), R = alpha c, A = alpha; loss is
    # smooth-L1 of R*A + bg(1-A) - gt_composite
    sigma, dt, c = 2.0, 0.3, 0.7
    s = np.array([math.log(sigma) + 1.0])
    alpha = 1 - math.exp(-sigma * dt)
    bg = np.array([0.2, 0.2, 0.2])
    target = np.array([0.5, 0.5, 0.5, 1.0])
    loss, _, _ = _comp([1], s, np.array([c, c, c]), np.array([dt]), bg, target)
    R = alpha * c
    pred = R * alpha + 0.2 * (1 - alpha)
    assert abs(loss[0] - 0.5 * (pred - 0.5) ** 2) < 1e-15
    assert abs(loss[3] - 0.5 * ((1 - alpha) - 0.0) ** 2) < 1e-15
    # constant sigma, no early stop: sum w + T_final = 1 and T_n = exp(-sigma sum dt)
    n = 20
    s = np.full(n, math.log(0.5) + 1.0)
    dts = np.full(n, 0.1)
    col = np.zeros(3 * n)
    loss, _, _ = _comp([n], s, col, dts, np.zeros(3), np.array([0, 0, 0, 0.0]), min_T=0.0)
    A = 1 - math.exp(-0.5 * 0.1 * n)
    # alpha channel loss = smoothL1((1-A) - 1)
    assert abs(loss[3] - 0.5 * A * A) < 1e-14


def test_composite_early_termination_index():
    # opaque samples: T drops below 1e-4 after k samples; later samples get no gradient
    s = np.full(10, math.log(10.0) + 1.0)
    dts = np.full(10, 0.3)  # T *= exp(-3) = 0.0498 per sample
    _, dLd, dLc = _comp([10], s, np.full(30, 0.5), dts, np.zeros(3), np.array([1, 1, 1, 1.0]))
    T, k = 1.0, 0
    while True:
        T *= math.exp(-3.0)
        k += 1
        if T < 1e-4:
            break
    assert np.all(dLd[:k] != 0) and np.all(dLd[k:] == 0)
    assert np.all(dLc[k:10] == 0)


def test_composite_backward_finite_differences():
    rng = np.random.default_rng(6)
    steps = np.clip(rng.geometric(1 / 12, 8), 1, 64)
    ns = int(steps.sum())
    s = np.log(rng.lognormal(0, 1.5, ns)) + 1.0
    c = rng.uniform(0, 1, 3 * ns)
    dt = rng.uniform(0.02, 0.2, ns)
    bg = rng.uniform(0, 1, 3 * 8)
    tg = rng.uniform(0, 1, 4 * 8)
    _, dLd, dLc = _comp(steps, s, c, dt, bg, tg)
    h = 1e-6

    def L(s_, c_):
        loss, _, _ = _comp(steps, s_, c_, dt, bg, tg)
        return loss.reshape(4, 8)[:3].sum() / 8

    for i in rng.choice(ns, 25, replace=False):
        e = np.zeros(ns); e[i] = h
        fd = (L(s + e, c) - L(s - e, c)) / (2 * h)
        assert abs(fd - dLd[i]) <= 1e-6 * max(1e-3, abs(fd))
        e3 = np.zeros(3 * ns); e3[i] = h
        fdc = (L(s, c + e3) - L(s, c - e3)) / (2 * h)
        assert abs(fdc - dLc[i]) <= 1e-6 * max(1e-3, abs(fdc))


# --------------------------------------------------------------- distortion
def test_distortion_linear_equals_quadratic_and_fd():
    rng = np.random.default_rng(7)
    for n in (1, 2, 5, 33, 64):
        d = np.log(rng.lognormal(0, 1.5, n)) + 1.0
        dt = rng.uniform(0.5, 1.5, n)
        t = np.concatenate([[0], np.cumsum(dt)])
        span = t[-1]
        m = ((t[:-1] + 0.5 * dt) / span)
        dtn = dt / span
        loss, grad = oracle.distortion(1, np.array([n], np.uint32), np.array([0], np.uint32), d, m, dtn, lam=0.01)
        bf = oracle.distortion_bruteforce(d, m, dtn, lam=0.01)
        assert abs(loss[0] - bf) <= 1e-12 * max(1.0, abs(bf))
        h = 1e-6
        for i in range(min(n, 8)):
            e = np.zeros(n); e[i] = h
            fd = (oracle.distortion_bruteforce(d + e, m, dtn, 0.01) - oracle.distortion_bruteforce(d - e, m, dtn, 0.01)) / (2 * h)
            assert abs(fd - grad[i]) <= 1e-6 * max(1e-6, abs(fd)) + 1e-12


# --------------------------------------------------------------- Adam
def test_adam_closed_form():
    n = 8
    w = np.linspace(-1, 1, n).astype(F32)
    g = np.array([0, 2048, -1024, 0, 512, 0, 256, -128], F32)  # loss scale 1024
    m = np.zeros(n, F32); v = np.zeros(n, F32); st = np.zeros(n, np.uint32)
    n_grid, n_net_end = 4, 7
    lr, b1, b2, eps, l2 = 1e-2, 0.9, 0.99, 1e-15, 1e-6
    w2, wl, m2, v2, st2 = oracle.adam_step(n_grid, n_net_end, 1024.0, lr, b1, b2, eps, l2, w, g, m, v, st)
    for i in range(n):
        gi = float(g[i]) / 1024
        if i < n_grid and gi == 0:
            assert w2[i] == w[i] and st2[i] == 0 and m2[i] == 0
            continue
        if n_grid <= i < n_net_end:
            gi += l2 * float(w[i])
        mi = (1 - b1) * gi
        vi = (1 - b2) * gi * gi
        step = lr * math.sqrt(1 - b2) / (1 - b1) * mi / (math.sqrt(vi) + eps)
        assert st2[i] == 1
        assert abs(float(w2[i]) - (float(w[i]) - step)) <= 1e-6
    assert w2[5] != w[5]  # network weight with zero grad still steps (L2 term)


# --------------------------------------------------------------- occupancy bits
def test_occupancy_bits_and_popcount():
    rng = np.random.default_rng(8)
    vol, L = 4096, 5
    dens = rng.uniform(0, 0.02, L * vol).astype(F32)
    bits = oracle.occ_update_bits(dens, vol, L, 0.01)
    expect = np.zeros(vol, np.uint8)
    for l in range(L):
        expect |= ((dens[l * vol:(l + 1) * vol] > F32(0.01)).astype(np.uint8) << l)
    assert np.array_equal(bits, expect)
    assert oracle.popcount(bits) == int(np.unpackbits(bits).sum())
    d2 = oracle.occ_decay(dens, L, 0.95)
    assert np.array_equal(d2, dens * F32(0.95))


def test_init_training_rays_camera_model_backwards():
    """Training rays pinned by running the camera model backwards in float64:
    the direction rotated into the camera frame and divided by its z gives the
    pixel's ((x + offset + 0.5) / fx, (y + offset + 0.5) / fy), x = fmod(p, W)
    and y = p / W left unfloored, p = n_pix * u; the origin lies on the near
    plane along that direction; t and t_max meet the box (slab test) or
    far - near; the target is texel floor(p) / 255 of image floor(i / rays per
    image); rays that miss the box are dead."""
    from turbonerf_amd import scenes

    NC, W, H, R = 6, 48, 40, 3000
    seed = 31
    rng = np.random.default_rng(seed)
    cams = scenes.sphere_cameras(seed, NC, 2.0, W, H, 50.0)
    images = rng.integers(0, 256, size=(NC, H, W, 4), dtype=np.uint8)
    rnd = rng.uniform(0, 1, R).astype(F32)
    rr = oracle.init_training_rays(R, R, NC, W, H, R / NC, scenes.bbox(1.0), cams, images, rnd)
    alive = rr["alive"].astype(bool)
    ori, d = rr["ori"].reshape(3, R).astype(np.float64), rr["dir"].reshape(3, R).astype(np.float64)
    img = (np.arange(R) * NC) // R
    assert np.array_equal(rr["img_id"][alive], img[alive])
    p = F32(W * H) * rnd
    x, y = np.fmod(p, F32(W)).astype(np.float64), (p / F32(W)).astype(np.float64)
    assert np.any(y[alive] != np.floor(y[alive]))
    # which rays hit the box: the pinhole model in float64 and a slab test
    hit = np.zeros(R, bool)
    clear = np.zeros(R, bool)
    for i in range(R):
        cam = cams[img[i]]
        M = np.array(cam.transform[:], np.float64).reshape(3, 4)
        v = np.array([(x[i] + cam.offset_px[0] + 0.5) / cam.focal_length[0],
                      (y[i] + cam.offset_px[1] + 0.5) / cam.focal_length[1], 1.0])
        dm = M[:, :3] @ v
        dm /= np.linalg.norm(dm)
        om = M[:, 3] + cam.near_plane * np.linalg.norm(v) * dm
        with np.errstate(divide="ignore"):
            ta, tb = (-0.5 - om) / dm, (0.5 - om) / dm
        a0, a1 = np.minimum(ta, tb).max(), np.maximum(ta, tb).min()
        hit[i] = a1 >= a0 and min(cam.far_plane - cam.near_plane, a1) >= max(0.0, a0 + 1e-5)
        clear[i] = abs(a1 - a0) > 1e-4
    assert np.array_equal(alive[clear], hit[clear])
    # t and t_max from a slab test on the oracle's own origin and direction
    with np.errstate(divide="ignore", invalid="ignore"):
        t_a, t_b = (-0.5 - ori) / d, (0.5 - ori) / d
    t0, t1 = np.minimum(t_a, t_b).max(0), np.maximum(t_a, t_b).min(0)
    checked = 0
    for i in np.nonzero(clear & alive)[0]:
        cam = cams[img[i]]
        M = np.array(cam.transform[:], np.float64).reshape(3, 4)
        checked += 1
        dc = M[:, :3].T @ d[:, i]
        want = np.array([(x[i] + cam.offset_px[0] + 0.5) / cam.focal_length[0],
                         (y[i] + cam.offset_px[1] + 0.5) / cam.focal_length[1]])
        np.testing.assert_allclose(dc[:2] / dc[2], want, rtol=1e-5, atol=2e-6)
        oc = M[:, :3].T @ (ori[:, i] - M[:, 3])
        assert abs(oc[2] - cam.near_plane) < 1e-6
        np.testing.assert_allclose(oc[:2] / oc[2], want, rtol=1e-4, atol=1e-5)
        assert abs(rr["t"][i] - max(0.0, t0[i] + 1e-5)) < 2e-6
        assert abs(rr["tmax"][i] - min(cam.far_plane - cam.near_plane, t1[i])) < 2e-6
        lp = min(int(np.floor(p[i])), W * H - 1)
        np.testing.assert_array_equal(rr["pix_rgba"].reshape(4, R)[:, i],
                                      images[img[i]].reshape(-1, 4)[lp].astype(F32) / F32(255))
    assert checked > R // 8 and (~alive).sum() > 0


# --------------------------------------------------------------- lens model
LENSES = [(-0.28, 0.07, 0.0, 1e-4, -2e-4),   # barrel, typical wide-angle calibration
          (0.12, -0.02, 0.003, -5e-4, 8e-4),   # pincushion with a cubic term
          (0.0, 0.0, 0.0, 2e-3, 0.0)]          # tangential only


def _lens_forward(x, y, k):
    """Radial-tangential (Brown-Conrady) distortion as OpenCV documents it,
    in float64: r = x^2 + y^2, radial factor 1 + k1 r + k2 r^2 + k3 r^3,
    tangential terms 2 p1 x y + p2 (r + 2 x^2) and p1 (r + 2 y^2) + 2 p2 x y."""
    k1, k2, k3, p1, p2 = (float(F32(v)) for v in k)
    r = x * x + y * y
    rad = 1.0 + k1 * r + k2 * r * r + k3 * r * r * r
    return x * rad + 2.0 * p1 * x * y + p2 * (r + 2.0 * x * x), y * rad + p1 * (r + 2.0 * y * y) + 2.0 * p2 * x * y


@pytest.mark.parametrize("k", LENSES)
def test_lens_residual_and_jacobian_against_forward_model(k):
    """The residual is the forward model minus the target and the Jacobian
    matches float64 central differences of the forward model."""
    rng = np.random.default_rng(41)
    h = 1e-6
    for x, y, xd, yd in rng.uniform(-0.6, 0.6, size=(200, 4)).astype(F32):
        f = oracle.lens_residual_jacobian(x, y, xd, yd, k)
        X, Y = float(x), float(y)
        fx, fy = _lens_forward(X, Y, k)
        np.testing.assert_allclose(f[:2], [fx - float(xd), fy - float(yd)], rtol=0, atol=3e-7)
        ax, ay = _lens_forward(X + h, Y, k)
        bx, by = _lens_forward(X - h, Y, k)
        cx, cy = _lens_forward(X, Y + h, k)
        ex, ey = _lens_forward(X, Y - h, k)
        jac = [(ax - bx) / (2 * h), (cx - ex) / (2 * h), (ay - by) / (2 * h), (cy - ey) / (2 * h)]
        np.testing.assert_allclose(f[2:], jac, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("k", LENSES)
def test_lens_undistort_inverts_forward_model(k):
    """Newton's answer, pushed through the forward model, lands on the
    distorted point to float32 rounding; with no distortion it is the point."""
    rng = np.random.default_rng(42)
    pts = rng.uniform(-0.55, 0.55, size=(400, 2)).astype(F32)
    worst = 0.0
    for xd, yd in pts:
        xu, yu = oracle.lens_undistort(xd, yd, k)
        gx, gy = _lens_forward(float(xu), float(yu), k)
        worst = max(worst, abs(gx - float(xd)), abs(gy - float(yd)))
        assert (xu, yu) != (xd, yd) or k[0] == 0.0
    assert worst < 3e-7, worst
    assert oracle.lens_undistort(F32(0.3), F32(-0.2), (0, 0, 0, 0, 0)) == (F32(0.3), F32(-0.2))


def test_distorted_camera_rays_model_backwards():
    """Training and render rays of lens-distorted cameras: the direction in
    the camera frame gives the undistorted pixel (u f - offset - 0.5 per
    axis), and that pixel, normalised about the principal point and pushed
    through the forward lens model, is the sampled pixel normalised the same
    way (training: x = fmod(p, W), y = p / W; render: pixel centres' corners
    ix, iy)."""
    from turbonerf_amd import scenes

    NC, W, H, R = 3, 64, 48, 2000
    rng = np.random.default_rng(43)
    images = rng.integers(0, 256, size=(NC, H, W, 4), dtype=np.uint8)
    rnd = rng.uniform(0, 1, R).astype(F32)
    img = (np.arange(R) * NC) // R
    p = F32(W * H) * rnd
    x, y = np.fmod(p, F32(W)).astype(np.float64), (p / F32(W)).astype(np.float64)

    def check(cam, d, px, py):
        M = np.array(cam.transform[:], np.float64).reshape(3, 4)
        dc = M[:, :3].T @ d
        f = np.array(cam.focal_length[:], np.float64)
        pp = np.array(cam.principal_point[:], np.float64)
        off = np.array(cam.offset_px[:], np.float64)
        u = (dc[:2] / dc[2] * f - off - 0.5 - pp) / f
        g = _lens_forward(u[0], u[1], cam.dist_params[:])
        np.testing.assert_allclose(g, [(px - pp[0]) / f[0], (py - pp[1]) / f[1]], rtol=0, atol=2e-5)
        return u

    for k in LENSES:
        cams = scenes.sphere_cameras(44, NC, 2.0, W, H, 60.0, dist_params=k)
        rr = oracle.init_training_rays(R, R, NC, W, H, R / NC, scenes.bbox(1.0), cams, images, rnd)
        alive = rr["alive"].astype(bool)
        d = rr["dir"].reshape(3, R).astype(np.float64)
        assert alive.sum() > R // 8
        moved = 0.0
        for i in np.nonzero(alive)[0]:
            u = check(cams[img[i]], d[:, i], x[i], y[i])
            moved = max(moved, abs(u[0] * cams[0].focal_length[0] + cams[0].principal_point[0] - x[i]))
        assert moved > 0.01  # pixels move 20x the tolerance: the pinhole model fails the check
        cam = cams[0]
        n = W * H
        rd = oracle.render_rays(n, 0, n, cam, scenes.bbox(1.0))
        dr = rd["dir"].reshape(3, n).astype(np.float64)
        for i in range(0, n, 37):
            check(cam, dr[:, i], float(i % W), float(i // W))


def _morton_decode(m):
    m = np.asarray(m, np.uint64)
    out = []
    for axis in range(3):
        v = np.zeros_like(m)
        for b in range(10):
            v |= ((m >> np.uint64(3 * b + axis)) & np.uint64(1)) << np.uint64(b)
        out.append(v.astype(np.int64))
    return out


def test_occ_sample_points_lie_in_their_cells():
    """Jittered occupancy points: for cell idx (Morton, decoded here by bit
    de-interleaving) of cascade k, the world point (pos - 0.5) * aabb minus the
    cell corner ((i - res/2) * 2^k / res per axis), over the cell size, is the
    uniform draw; positions are not clamped to [0, 1]."""
    res, n, start, level, aabb = 128, 4096, 777777, 3, 4.0
    rng = np.random.default_rng(9)
    rnd = rng.uniform(0, 1, 3 * n).astype(F32)
    pos = oracle.occ_sample_points(n, n, start, res, level, 1.0 / aabb, rnd).reshape(3, n).astype(np.float64)
    idx = np.arange(start, start + n)
    ijk = _morton_decode(idx)
    voxel = 2.0 ** level / res
    for a in range(3):
        corner = (ijk[a] - res // 2) * voxel
        u = ((pos[a] - 0.5) * aabb - corner) / voxel
        np.testing.assert_allclose(u, rnd[a * n:(a + 1) * n], atol=2e-5)
    assert np.any(pos < 0.0) or np.any(pos > 1.0)


def test_occ_update_density_subset_rule():
    """Known answers at the rule's boundaries: after warm-up a cell is updated
    when r < 0.25, or r <= 0.5 and its bit at this cascade is set, never when
    r > 0.5; during warm-up always; the update is max(old, exp(min(s, 12) - 1))
    at this cascade only."""
    vol, L, level, start, n = 4096, 3, 1, 100, 2000
    rng = np.random.default_rng(10)
    r = rng.uniform(0, 1, n).astype(F32)
    r[:8] = [0.25, 0.5, np.nextafter(F32(0.5), F32(1)), np.nextafter(F32(0.25), F32(0)), 0.0, 1.0, 0.5, 0.25]
    bits = rng.integers(0, 8, vol).astype(np.uint8)
    bit = np.uint8(1 << level)
    for k, on in ((0, False), (1, True), (2, True), (6, False), (7, True)):  # boundary cases, both ways
        bits[start + k] = (bits[start + k] | bit) if on else (bits[start + k] & ~bit)
    raw = rng.uniform(-3, 15, n).astype(F32)
    dens = rng.uniform(0, 20, L * vol).astype(F32)
    sig = np.exp(np.minimum(raw.astype(np.float64), 12.0) - 1.0)
    occ = ((bits[start:start + n] >> level) & 1).astype(bool)
    for warm in (True, False):
        got = oracle.occ_update_density(n, start, vol, level, warm, r, raw, bits, dens)
        sel = np.ones(n, bool) if warm else (r < 0.25) | ((r <= 0.5) & occ)
        if not warm:
            assert list(sel[:8]) == [False, True, False, True, True, False, False, True]
        old = dens[level * vol + start: level * vol + start + n].astype(np.float64)
        want = np.where(sel, np.maximum(old, sig), old)
        np.testing.assert_allclose(got[level * vol + start: level * vol + start + n], want, rtol=1e-6)
        keep = np.ones(L * vol, bool)
        keep[level * vol + start: level * vol + start + n] = False
        assert np.array_equal(got[keep], dens[keep])


# --------------------------------------------------------------- hash grid
def _layout():
    from turbonerf_amd import _capi

    g = _capi.HashGrid()
    b = math.exp((math.log(2048) - math.log(16)) / 15)
    oracle.lib().oc_hashgrid_layout(16, 2, 19, 16, b, ctypes.addressof(g))
    return g


def _paper_layout(L, log2_T, n_min, b, vertex_rule="paper"):
    """Level table from the Instant-NGP paper (Mueller et al. 2022, sec. 3):
    N_l = floor(N_min * b^l) cells per axis (eq. 2), hence N_l + 1 vertices;
    a level whose dense grid fits the table, (N_l + 1)^3 <= T, is indexed 1:1,
    otherwise through the spatial hash. Levels hold min(vertices^3 rounded up
    to 8, T) entries. vertex_rule="paper-1" and "strict" are planted mistakes
    (one vertex too few; a strict dense threshold) that the test shows the
    known-answer table rejects."""
    T = 1 << log2_T
    rows = []
    off = 0
    for l in range(L):
        x = n_min * b ** l
        v = math.floor(x) + 1
        if vertex_rule == "paper-1":
            v -= 1
        vol = v ** 3
        dense = vol < T if vertex_rule == "strict" else vol <= T
        size = min((vol + 7) // 8 * 8, T)
        rows.append((v, dense, size, off, x))
        off += size
    return rows, 2 * off


def test_hashgrid_layout_matches_instant_ngp_equations():
    """The level table at the training configuration (16 levels, T = 2^19,
    N 16 -> 2048) against the paper's equations. tiny-cuda-nn, which the
    reference configures (src/core/nerf-network.cu:86-104), counts
    ceil(N_min b^l - 1) + 1 vertices per axis; that equals the paper's
    floor(N_min b^l) + 1 wherever N_min b^l is not an integer and is one
    fewer where it is (level 0 here: 16 vertices where the paper has 17).
    The build keeps tiny-cuda-nn's count; this test fixes both the agreement
    and that single departure, the dense rule (N_l + 1)^3 <= T, the level
    sizes and offsets, and the known-answer parameter count 12,196,240."""
    from turbonerf_amd import _capi, ops

    b = math.exp((math.log(2048) - math.log(16)) / 15)
    rows, n_params = _paper_layout(16, 19, 16, b)
    # known answers of the paper's equations, written out
    assert [r[0] for r in rows[:6]] == [17, 23, 31, 43, 59, 81]
    assert [r[1] for r in rows] == [True] * 5 + [False] * 11
    for name, g in (("oracle", _layout()), ("product", ops.hashgrid_layout())):
        for l, (v, dense, size, off, x) in enumerate(rows):
            integral = abs(x - round(x)) < 1e-9
            want_v = v - 1 if integral else v
            assert g.resolution[l] == want_v, (name, l, g.resolution[l], want_v)
            assert bool(g.level_dense[l]) == (want_v ** 3 <= (1 << 19)), (name, l)
            if l > 0:
                assert g.level_size[l] == size and g.level_dense[l] == dense, (name, l)
            # position scale: x * (vertices - 1) spans the level's vertices
            assert abs(g.scale[l] - np.float32(x - 1)) <= 1e-6 * x, (name, l)
        assert g.level_size[0] == 16 ** 3 and g.level_offset[1] == 16 ** 3
        assert [g.level_offset[l] for l in range(16)] == list(
            np.cumsum([0] + [g.level_size[l] for l in range(15)])), name
        assert g.n_params == 12_196_240, (name, g.n_params)
    # the paper's table differs from the build's only by level 0's one vertex
    assert n_params == 12_196_240 + 2 * (4920 - 4096) == 12_197_888
    # planted mistakes in the paper restatement are caught by the known answers
    bad, _ = _paper_layout(16, 19, 16, b, "paper-1")
    assert [r[0] for r in bad[:6]] != [17, 23, 31, 43, 59, 81]
    # dense-threshold boundary: a level with exactly T vertices is dense under
    # "<=" and hashed under "<"; the build follows "<="
    b2 = 63.5 / 16.0
    rows2, _ = _paper_layout(2, 18, 16, b2)
    strict2, _ = _paper_layout(2, 18, 16, b2, "strict")
    assert rows2[1][0] == 64 and rows2[1][1] and not strict2[1][1]
    g2 = _capi.HashGrid()
    oracle.lib().oc_hashgrid_layout(2, 2, 18, 16, b2, ctypes.addressof(g2))
    g3 = ops.hashgrid_layout(2, 2, 18, 16, per_level_scale=b2)
    for g in (g2, g3):
        assert g.resolution[1] == 64 and g.level_dense[1] == 1 and g.level_size[1] == 1 << 18
    oracle.lib().oc_hashgrid_layout(2, 2, 17, 16, b2, ctypes.addressof(g2))
    assert g2.level_dense[1] == 0 and g2.level_size[1] == 1 << 17


def test_hashgrid_partition_of_unity_and_adjoint():
    g = _layout()
    rng = np.random.default_rng(9)
    n = 300
    pos = rng.uniform(0, 1, 3 * n).astype(F32)
    ones = np.ones(g.n_params)
    feat = np.zeros(32 * n)
    oracle.lib().oc_hashgrid_fwd_f64(ctypes.addressof(g), n, n, pos, ones, feat)
    np.testing.assert_allclose(feat, 1.0, atol=1e-12)
    table = rng.normal(size=g.n_params)
    oracle.lib().oc_hashgrid_fwd_f64(ctypes.addressof(g), n, n, pos, table, feat)
    G = rng.normal(size=32 * n)
    dt = np.zeros(g.n_params)
    oracle.lib().oc_hashgrid_bwd_f64(ctypes.addressof(g), n, n, pos, G, dt)
    assert abs(float(feat @ G) - float(table @ dt)) <= 1e-9 * abs(float(feat @ G))


def test_hashgrid_index_formula_bruteforce():
    g = _layout()
    # level 0 is dense: a point inside cell (i,j,k) with frac 0 reads entry i + j r + k r^2
    r = g.resolution[0]
    table = np.zeros(g.n_params)
    i, j, k = 3, 5, 7
    e = i + j * r + k * r * r
    table[2 * e] = 1.0
    s = g.scale[0]
    p = np.array([(i - 0.5) / s, (j - 0.5) / s, (k - 0.5) / s], F32)
    feat = np.zeros(32)
    oracle.lib().oc_hashgrid_fwd_f64(ctypes.addressof(g), 1, 1, p, table, feat)
    x = p * F32(s) + F32(0.5)
    w = np.prod([1 - (xx - np.floor(xx)) for xx in x])
    assert abs(feat[0] - w) < 1e-12
    # hashed level: the XOR-of-primes formula
    l = 10
    assert not g.level_dense[l]
    cell = (100, 200, 300)
    h = (cell[0] * 1) ^ ((cell[1] * 2654435761) & 0xFFFFFFFF) ^ ((cell[2] * 805459861) & 0xFFFFFFFF)
    ent = g.level_offset[l] + (h & 0xFFFFFFFF) % g.level_size[l]
    table = np.zeros(g.n_params)
    table[2 * ent + 1] = 1.0
    sl = g.scale[l]
    p = ((np.array(cell, np.float64) - 0.5 + 1e-3) / sl).astype(F32)
    feat = np.zeros(32)
    oracle.lib().oc_hashgrid_fwd_f64(ctypes.addressof(g), 1, 1, p, table, feat)
    assert feat[2 * l + 1] > 0.9


# --------------------------------------------------------------- spherical harmonics
def test_sh_orthonormal_quadrature():
    # Gauss-Legendre in cos(theta) x uniform phi integrates products of two
    # degree-4 harmonics (degree 8 polynomials) exactly
    xs, ws = np.polynomial.legendre.leggauss(10)
    phis = np.arange(20) * 2 * np.pi / 20
    z = np.repeat(xs, 20)
    wq = np.repeat(ws, 20) * (2 * np.pi / 20)
    ph = np.tile(phis, 10)
    r = np.sqrt(1 - z * z)
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    dir01 = (0.5 * d + 0.5).astype(F32)
    n = len(z)
    out = np.zeros(16 * n)
    oracle.lib().oc_sh4_f64(n, n, np.ascontiguousarray(dir01.T.ravel()), out)
    Y = out.reshape(16, n)
    gram = (Y * wq) @ Y.T
    np.testing.assert_allclose(gram, np.eye(16), atol=2e-6)  # inputs are fp32 directions


# --------------------------------------------------------------- MLP reference
def test_numpy_mlp_backward_matches_autograd():
    import torch

    from nets import backward, forward, random_net, sh4

    rng = np.random.default_rng(11)
    n = 40
    net = random_net(3)
    feat = rng.normal(size=(n, 32))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir01 = 0.5 * d + 0.5
    G = rng.normal(size=(n, 3))
    dls = rng.normal(size=n)
    t = rng.uniform(0.2, 2.0, n)
    fw = forward(net, feat, dir01)
    dX, gr = backward(net, fw, G, dls, t, 7.0)
    W = torch.tensor(net, dtype=torch.float64, requires_grad=True)
    X = torch.tensor(feat, dtype=torch.float64, requires_grad=True)
    W1, W2, W3, W4, W5 = W[:2048].view(64, 32), W[2048:3072].view(16, 64), W[3072:5120].view(64, 32), \
        W[5120:9216].view(64, 64), W[9216:].view(3, 64)
    h1 = torch.relu(X @ W1.T)
    s16 = h1 @ W2.T
    # the t^2 factor scales the gradient reaching s16[:, 0]; express it as a
    # hook on that column
    scale = torch.tensor(np.where(t < 1, t * t, 1.0))
    s16 = s16.clone()
    col0 = s16[:, 0]
    s16h = torch.cat([(col0 * 1.0).unsqueeze(1), s16[:, 1:]], 1)
    s16h.register_hook(lambda gr_: torch.cat([(gr_[:, 0] * scale).unsqueeze(1), gr_[:, 1:]], 1))
    xc = torch.cat([s16h, torch.tensor(sh4(dir01))], 1)
    rgb = torch.sigmoid(torch.relu(torch.relu(xc @ W3.T) @ W4.T) @ W5.T)
    total = (rgb * torch.tensor(G)).sum() + 7.0 * (s16h[:, 0] * torch.tensor(dls)).sum()
    total.backward()
    np.testing.assert_allclose(gr, W.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(dX, X.grad.numpy(), rtol=1e-10, atol=1e-12)


def test_numpy_appearance_embedding_matches_autograd():
    """The embedding's concatenated input (reference nerf-network.cu:494-520)
    and its gradients in the numpy restatement against torch autograd."""
    import torch

    from nets import backward, forward, random_net, sh4

    rng = np.random.default_rng(13)
    n, n_img = 48, 5
    net = random_net(4)
    w3e = rng.uniform(-0.2, 0.2, size=(64, 16))
    emb = rng.uniform(0.0, 1.0, size=(n_img, 16))
    img = np.sort(rng.integers(0, n_img - 1, n))  # image 4 unused: its gradient is zero
    feat = rng.normal(size=(n, 32))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir01 = 0.5 * d + 0.5
    G = rng.normal(size=(n, 3))
    dls = rng.normal(size=n)
    t = rng.uniform(1.0, 2.0, n)
    fw = forward(net, feat, dir01, app=(w3e, emb, img))
    dX, gr, g_w3e, g_emb = backward(net, fw, G, dls, t, 3.0)
    W = torch.tensor(net, dtype=torch.float64, requires_grad=True)
    We = torch.tensor(w3e, requires_grad=True)
    E = torch.tensor(emb, requires_grad=True)
    X = torch.tensor(feat, requires_grad=True)
    W1, W2, W3, W4, W5 = W[:2048].view(64, 32), W[2048:3072].view(16, 64), W[3072:5120].view(64, 32), \
        W[5120:9216].view(64, 64), W[9216:].view(3, 64)
    s16 = torch.relu(X @ W1.T) @ W2.T
    xc = torch.cat([s16, torch.tensor(sh4(dir01)), E[torch.tensor(img)]], 1)
    rgb = torch.sigmoid(torch.relu(torch.relu(xc @ torch.cat([W3, We], 1).T) @ W4.T) @ W5.T)
    total = (rgb * torch.tensor(G)).sum() + 3.0 * (s16[:, 0] * torch.tensor(dls)).sum()
    total.backward()
    np.testing.assert_allclose(gr, W.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(g_w3e, We.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(g_emb, E.grad.numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(dX, X.grad.numpy(), rtol=1e-10, atol=1e-12)
    assert not g_emb[4].any() and np.abs(g_emb[:4]).min() >= 0 and np.abs(g_emb).max() > 0


def test_oracle_mlp_train_matches_numpy_reference():
    """The C MLP forward+backward (CPU-baseline training step) against the
    autograd-pinned numpy restatement: fp64 to 1e-12, fp32 to 1e-4."""
    from nets import backward, forward, random_net

    rng = np.random.default_rng(12)
    n = 300
    net = random_net(5)
    feat = rng.normal(scale=0.5, size=(n, 32))
    d = rng.normal(size=(n, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir01 = (0.5 * d + 0.5).astype(np.float32)
    G = rng.normal(size=(n, 3))
    dls = rng.normal(scale=1e-2, size=n).astype(np.float32)
    ex = rng.normal(scale=1e-3, size=n).astype(np.float32)
    t = rng.uniform(0.2, 2.0, n).astype(np.float32)
    fw = forward(net, feat, dir01)
    dX, gr = backward(net, fw, G, dls, t, 5.0, ex)
    dens, rgb, dfeat, dW = oracle.mlp_train(net, feat.T, dir01.T, G.T, dls, t, 5.0, extra=ex)
    np.testing.assert_allclose(rgb.T, fw["rgb"], rtol=1e-12)
    np.testing.assert_allclose(dens, fw["s16"][:, 0], rtol=1e-12, atol=1e-14)
    dens_f, rgb_f, _, dW_f = oracle.mlp_train(net, feat.T, dir01.T, None, None, None, 5.0)
    assert np.array_equal(dens_f, dens) and np.array_equal(rgb_f, rgb) and not dW_f.any()
    np.testing.assert_allclose(dW, gr, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(dfeat.T, dX, rtol=1e-9, atol=1e-12)
    _, _, dfeat32, dW32 = oracle.mlp_train(net, feat.T, dir01.T, G.T, dls, t, 5.0, extra=ex, fp64=False)
    assert np.linalg.norm(dW32 - gr) / np.linalg.norm(gr) < 1e-4
    assert np.linalg.norm(dfeat32.T - dX) / np.linalg.norm(dX) < 1e-4


def test_oracle_hashgrid_f32_matches_f64():
    import ctypes as C

    g = _capi_hashgrid()
    rng = np.random.default_rng(13)
    n = 2000
    pos = rng.uniform(0, 1, size=3 * n).astype(np.float32)
    table = rng.uniform(-1, 1, size=g.n_params).astype(np.float32)
    f64 = np.zeros(32 * n)
    f32 = np.zeros(32 * n, np.float32)
    oracle.lib().oc_hashgrid_fwd_f64(C.addressof(g), n, n, pos, table.astype(np.float64), f64)
    oracle.lib().oc_hashgrid_fwd_f32(C.addressof(g), n, n, pos, table, f32)
    np.testing.assert_allclose(f32, f64, rtol=1e-5, atol=1e-6)
    dfeat = rng.normal(size=32 * n)
    b64 = np.zeros(g.n_params)
    b32 = np.zeros(g.n_params, np.float32)
    oracle.lib().oc_hashgrid_bwd_f64(C.addressof(g), n, n, pos, dfeat, b64)
    oracle.lib().oc_hashgrid_bwd_f32(C.addressof(g), n, n, pos, dfeat.astype(np.float32), b32)
    assert np.linalg.norm(b32 - b64) / np.linalg.norm(b64) < 1e-6


def _capi_hashgrid():
    import ctypes as C
    import math

    from turbonerf_amd import _capi

    g = _capi.HashGrid()
    b = math.exp((math.log(2048.0) - math.log(16.0)) / 15.0)
    oracle.lib().oc_hashgrid_layout(16, 2, 19, 16, b, C.addressof(g))
    return g


# --------------------------------------------------------------- C ABI surface
def test_c_abi_exports_every_declared_symbol():
    import re
    from pathlib import Path

    from turbonerf_amd import _capi

    hdr = (Path(__file__).resolve().parent.parent / "include" / "turbonerf" / "tn_c_api.h").read_text()
    declared = set(re.findall(r"^(?:int|size_t|uint32_t)\s+(tn_\w+)\(", hdr, re.M))
    assert len(declared) >= 20
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding covers all of them
    assert declared <= set(_capi.declared_symbols()), declared - set(_capi.declared_symbols())


def test_tuning_knobs_validate_and_count_changes():
    """Launch-shape knobs: defaults, refused out-of-range values (nothing
    changes), and the change counter a captured graph is keyed on (host-only
    entry points: no GPU needed)."""
    from turbonerf_amd import _capi

    lib = _capi.lib()
    v = ctypes.c_int()
    defaults = {0: 256, 1: 128, 2: 3, 3: 1, 4: 2048, 5: 256, 6: 1, 7: 0, 8: 0, 9: 0}
    for k, d in defaults.items():
        assert lib.tn_tune_get(k, ctypes.byref(v)) == 0 and v.value == d
    assert lib.tn_tune_get(10, ctypes.byref(v)) == 1 and lib.tn_tune_get(-1, ctypes.byref(v)) == 1
    g0 = lib.tn_tune_generation()
    bad = {0: [32, 192, 512], 1: [7, 257], 2: [0, 5], 3: [0, 3], 4: [255, 8193], 5: [0, 96, 1088], 6: [0, 17],
           7: [-1, 5], 8: [-1, 2], 9: [-1, 2], 10: [1]}
    for k, vals in bad.items():
        for x in vals:
            assert lib.tn_tune_set(k, x) == 1, (k, x)
    assert lib.tn_tune_generation() == g0
    assert lib.tn_tune_set(4, 2048) == 0 and lib.tn_tune_generation() == g0  # unchanged value
    assert lib.tn_tune_set(4, 1024) == 0 and lib.tn_tune_generation() == g0 + 1
    assert lib.tn_tune_get(4, ctypes.byref(v)) == 0 and v.value == 1024
    assert lib.tn_tune_set(4, 2048) == 0 and lib.tn_tune_generation() == g0 + 2
