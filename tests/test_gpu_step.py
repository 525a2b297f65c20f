"""One full fp32 training step through the C ABI, stage by stage against the
oracle on the same inputs (each stage fed the GPU's own previous output):
uniforms, rays, marching counts/offsets and every per-sample buffer
bit-exact; hash grid, MLP forward, compositing, MLP backward and hash-grid
scatter-add within fp32 rounding of the fp64 oracle (norm-wise 1e-5; compositing
1e-4); the
Adam update within 1e-6 of the oracle's on the same gradients. Rays are drawn
through random RGBA8 images of 16 cameras on a radius-2 sphere; 512 rays, a
budget of 2^13 samples; the occupancy grid is the thresholded blob field."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def norm_rel(a, b):
    a, b = np.ravel(np.asarray(a, np.float64)), np.ravel(np.asarray(b, np.float64))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_training_step_stages_vs_oracle(cuda):
    import torch

    import oracle
    from nets import random_net
    from turbonerf_amd import ops, scenes

    NC, W, H, R, B = 16, 64, 64, 512, 1 << 13
    seed = 77
    rng = np.random.default_rng(seed)
    cams = scenes.sphere_cameras(seed, NC, 2.0, W, H, 50.0)
    images = rng.integers(0, 256, size=(NC, H, W, 4), dtype=np.uint8)
    bbox = scenes.bbox(1.0)
    dt = float(np.float32(math.sqrt(3) / 1024))
    blobs = scenes.blob_field(seed, 8, 1.0)
    ii = np.arange(128 ** 3, dtype=np.uint32)
    inv = lambda m: sum(((m >> (3 * b)) & 1) << b for b in range(10))
    xyz = (np.stack([inv(ii), inv(ii >> 1), inv(ii >> 2)], 1).astype(np.float32) - 64 + 0.5) / 128
    sig, _ = oracle.blob_eval(xyz, blobs)
    bits = oracle.occ_update_bits(sig, 128 ** 3, 1, 0.01)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()

    # uniforms and rays
    rnd = ops.fill_uniform(4 * R, seed, 0)
    rnd_o = oracle.uniform(4 * R, seed, 0)
    assert np.array_equal(rnd.cpu().numpy(), rnd_o)
    rays = ops.Rays.empty(R)
    ops.init_training_rays(rays, R, NC, W, H, R / NC, ops.upload_struct(bbox), T(scenes.cameras_bytes(cams)),
                           T(images.ravel()), rnd)
    ro = oracle.init_training_rays(R, R, NC, W, H, R / NC, bbox, cams, images, rnd_o)
    for k, gk in (("ori", rays.ori), ("dir", rays.dir), ("t", rays.t), ("tmax", rays.tmax)):
        alive = ro["alive"].astype(bool)
        got = gk.cpu().numpy().reshape(-1, R)[:, alive]
        assert np.array_equal(got, ro[k].reshape(-1, R)[:, alive]), k
    assert np.array_equal(rays.alive.cpu().numpy(), ro["alive"])
    # targets of every ray, zero for rays that miss the box
    assert np.array_equal(rays.pix_rgba.cpu().numpy(), ro["pix_rgba"])

    # marching, scan, generation
    grid = ops.grid_struct(T(bits), None, 1, 128)
    ops.march_count(rays, R, bbox, grid, 1 / 256, dt, dt)
    n_steps_o, alive_o = oracle.march_count(R, R, bbox, bits, 1, 128, 1 / 256, dt, dt, ro)
    assert np.array_equal(rays.n_steps.cpu().numpy().view(np.uint32), n_steps_o)
    info = torch.zeros(4, dtype=torch.int32, device="cuda")
    ops.scan_select(rays, R, B, info)
    offs_o, nr, ns, rc = oracle.scan_select(n_steps_o, R, B)
    assert rc == 0 and info[0].item() == nr and info[1].item() == ns
    assert np.array_equal(rays.offset.cpu().numpy().view(np.uint32)[:nr], offs_o[:nr])
    smp = ops.Samples.empty(B)
    ops.march_generate(rays, smp, R, info, bbox, 1.0, grid, dt, dt, 1 / 256)
    so = oracle.march_generate(nr, R, B, bbox, 1.0, bits, 1, 128, dt, dt, 1 / 256, ro, offs_o, alive_o, n_steps_o)
    for k, gk in (("t", smp.t), ("pos", smp.pos), ("dir", smp.dir), ("dt", smp.dt), ("m_norm", smp.m_norm),
                  ("dt_norm", smp.dt_norm)):
        assert np.array_equal(gk.cpu().numpy().reshape(-1, B)[:, :ns], so[k].reshape(-1, B)[:, :ns]), k
    print(f"[step] rays {nr} samples {ns}", flush=True)

    # network forward
    g = ops.hashgrid_layout()
    table = rng.uniform(-0.5, 0.5, g.n_params).astype(np.float32)
    net = random_net(seed).astype(np.float32)
    params = T(np.concatenate([table, net]))
    feat = ops.hashgrid_forward(g, B, smp.pos, B, params, info=info)
    pos_c = np.ascontiguousarray(smp.pos.cpu().numpy().reshape(3, B)[:, :ns])
    f_ref = np.zeros(32 * ns)
    oracle.lib().oc_hashgrid_fwd_f64(ctypes.addressof(g), ns, ns, pos_c.ravel(), table.astype(np.float64), f_ref)
    feat_c = feat.cpu().numpy().reshape(32, B)[:, :ns]
    assert norm_rel(feat_c, f_ref) < 1e-5
    dens, rgb = ops.mlp_forward(params, g.n_params, B, feat, B, smp.dir, B, color=True, info=info)
    dir_c = np.ascontiguousarray(smp.dir.cpu().numpy().reshape(3, B)[:, :ns])
    d_ref, rgb_ref, _, _ = oracle.mlp_train(net.astype(np.float64), feat_c.astype(np.float64), dir_c, None, None,
                                            None, 1.0)
    dens_c, rgb_c = dens.cpu().numpy()[:ns], rgb.cpu().numpy().reshape(3, B)[:, :ns]
    assert norm_rel(dens_c, d_ref) < 1e-5 and norm_rel(rgb_c, rgb_ref) < 1e-5

    # compositing + loss + backward
    loss, dLd, dLc = ops.composite_train(info, R, R, B, rays.n_steps, rays.offset, rnd[R:], rays.pix_rgba, dens, rgb,
                                         smp.dt)
    bg = rnd_o[R:]
    dt_c = smp.dt.cpu().numpy()
    l_ref, dLd_ref, dLc_ref = oracle.composite_train(nr, R, B, n_steps_o, offs_o, bg, ro["pix_rgba"],
                                                     np.pad(dens_c, (0, B - ns)).astype(np.float64),
                                                     np.pad(rgb_c, ((0, 0), (0, B - ns))).ravel().astype(np.float64),
                                                     dt_c.astype(np.float64))
    assert norm_rel(loss.cpu().numpy().reshape(4, R)[:, :nr], l_ref.reshape(4, R)[:, :nr]) < 1e-5
    dLd_c = dLd.cpu().numpy()[:ns]
    dLc_c = dLc.cpu().numpy().reshape(3, B)[:, :ns]
    e_d, e_c = norm_rel(dLd_c, dLd_ref[:ns]), norm_rel(dLc_c, dLc_ref.reshape(3, B)[:, :ns])
    print(f"[step] composite dL/ds {e_d:.2e} dL/dc {e_c:.2e}", flush=True)
    # dense rays (table values up to 0.5) take transmittance near the
    # early-termination threshold, where fp32 products lose a few bits
    assert e_d < 1e-4 and e_c < 1e-4
    print(f"[step] loss {loss.cpu().numpy().reshape(4, R)[:, :nr].sum(0).mean():.5f}", flush=True)

    # network backward
    grads = torch.zeros_like(params)
    dfeat = ops.mlp_backward(params, g.n_params, B, feat, B, smp.dir, B, smp.t, dLc, dLd, 1.0, grads[g.n_params:],
                             info=info)
    t_c = np.ascontiguousarray(smp.t.cpu().numpy()[:ns])
    _, _, dfeat_ref, dW_ref = oracle.mlp_train(net.astype(np.float64), feat_c.astype(np.float64), dir_c,
                                               dLc_c.astype(np.float64), dLd_c, t_c, 1.0)
    dfeat_c = dfeat.cpu().numpy().reshape(32, B)[:, :ns]
    e_w, e_x = norm_rel(grads[g.n_params:].cpu().numpy(), dW_ref), norm_rel(dfeat_c, dfeat_ref)
    print(f"[step] mlp bwd dW {e_w:.2e} dfeat {e_x:.2e}", flush=True)
    assert e_w < 1e-5 and e_x < 1e-5
    ops.hashgrid_backward(g, B, smp.pos, B, dfeat, grads, info=info)
    dtab_ref = np.zeros(g.n_params)
    oracle.lib().oc_hashgrid_bwd_f64(ctypes.addressof(g), ns, ns, pos_c.ravel(),
                                     np.ascontiguousarray(dfeat_c, np.float64).ravel(), dtab_ref)
    grads_c = grads.cpu().numpy()
    e_t = norm_rel(grads_c[:g.n_params], dtab_ref)
    print(f"[step] hash bwd {e_t:.2e}", flush=True)
    assert e_t < 1e-5

    # Adam on the GPU's gradients
    P = params.numel()
    m1 = torch.zeros_like(params)
    m2 = torch.zeros_like(params)
    st = torch.zeros(P, dtype=torch.int32, device="cuda")
    w0 = params.cpu().numpy()
    args = (g.n_params, P, 1.0, 1e-2, 0.9, 0.99, 1e-15, 1e-6)
    ops.adam_step(*args, params, None, grads, m1, m2, st)
    w_ref, _, _, _, s_ref = oracle.adam_step(*args, w0, grads_c, np.zeros(P, np.float32), np.zeros(P, np.float32),
                                             np.zeros(P, np.uint32))
    w1 = params.cpu().numpy()
    np.testing.assert_allclose(w1, w_ref, rtol=1e-6, atol=1e-9)
    assert np.array_equal(st.cpu().numpy().view(np.uint32), s_ref)
    moved = np.count_nonzero(w1 != w0)
    print(f"[step] adam moved {moved} of {P} parameters", flush=True)
    assert moved > 9000
