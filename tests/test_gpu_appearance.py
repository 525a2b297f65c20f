"""Appearance embedding (reference src/core/embedding.cuh:45-99,
src/core/nerf-network.cu:27-72,494-520,698-715): the colour MLP's first layer
takes [density outputs | SH | emb[image]] through [W3 | W3e].

* Kernels: the per-image bias rows, the fused MLPs with them, and the W3e /
  embedding gradients from the colour layer's dz, against the fp64 numpy
  restatement that concatenates the embedding (pinned against torch
  autograd in test_oracle_pins.py). fp32 MFMA path: norm-wise 1e-5; fp16
  path: the forward within 2e-3 as without the embedding, the W3e and
  embedding gradients (which come from fp32 dz sums) within 1e-3, the
  weights within the bars of the network test. Samples sorted by image with
  one image absent (its gradient stays zero).
* Product: a Trainer with the embedding (one per training image) trains, its
  embeddings of the images in the batches move and the loss falls; a render
  uses image 0's embedding (changing image 1's leaves it unchanged, changing
  image 0's changes it); a snapshot of it stores 8192 colour parameters and
  16 per image, and loads back to the same render bit for bit.
"""
import numpy as np
import pytest

from nets import backward as np_backward
from nets import forward as np_forward
from nets import random_net

pytestmark = pytest.mark.gpu


def norm_rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _case(seed, n, n_img, dtype):
    import torch

    rng = np.random.default_rng(seed)
    net = random_net(seed)
    w3e = rng.uniform(-0.25, 0.25, size=(64, 16))
    emb = rng.uniform(0.0, 1.0, size=(n_img, 16))
    if dtype == torch.float16:
        net = net.astype(np.float16).astype(np.float64)
    w3e = w3e.astype(np.float32).astype(np.float64)
    emb = emb.astype(np.float32).astype(np.float64)
    img = np.sort(rng.integers(0, n_img, n))
    img[img == 2] = 3  # image 2 has no samples
    feat = rng.normal(scale=0.5, size=(n, 32)).astype(np.float32).astype(np.float64)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir01 = (0.5 * d + 0.5).astype(np.float32)
    return rng, net, w3e, emb, np.sort(img), feat, dir01


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_appearance_kernels_vs_fp64(cuda, prec):
    import torch

    from turbonerf_amd import ops

    dtype = torch.float32 if prec == "fp32" else torch.float16
    n, n_img = 3000, 6
    rng, net, w3e, emb, img, feat, dir01 = _case(21, n, n_img, dtype)
    app = (w3e, emb, img)
    n_grid = 64
    params = torch.zeros(n_grid + net.size, dtype=dtype, device="cuda")
    params[n_grid:] = torch.from_numpy(net).to(dtype).cuda()
    feat_d = torch.from_numpy(np.ascontiguousarray(feat.T.ravel())).float().cuda()
    dir_d = torch.from_numpy(np.ascontiguousarray(dir01.T.ravel())).cuda()
    w3e_d = torch.from_numpy(w3e.ravel()).float().cuda()
    emb_d = torch.from_numpy(emb.ravel()).float().cuda()
    img_d = torch.from_numpy(img.astype(np.int32)).cuda()
    bias = ops.appearance_bias(w3e_d, emb_d)
    np.testing.assert_allclose(bias.cpu().numpy().reshape(n_img, 64), emb @ w3e.T, rtol=1e-5, atol=1e-6)

    # forward
    dens, rgb = ops.mlp_forward(params, n_grid, n, feat_d, n, dir_d, n, color=True,
                                app=ops.appearance(bias, img_d))
    fw = np_forward(net, feat, dir01, app=app)
    tol = 1e-5 if prec == "fp32" else 2e-3
    e_rgb = norm_rel(rgb.float().cpu().numpy().reshape(3, n).T, fw["rgb"])
    e_d = norm_rel(dens.float().cpu().numpy(), fw["s16"][:, 0])
    print(f"[appearance fwd {prec}] rgb {e_rgb:.2e} density {e_d:.2e}", flush=True)
    assert e_rgb < tol and e_d < tol
    # the embedding moves the colours (a forward pass without it differs)
    _, rgb0 = ops.mlp_forward(params, n_grid, n, feat_d, n, dir_d, n, color=True)
    assert norm_rel(rgb0.float().cpu().numpy().reshape(3, n).T, fw["rgb"]) > 10 * tol
    # without image ids every sample takes row 0
    _, rgb_r0 = ops.mlp_forward(params, n_grid, n, feat_d, n, dir_d, n, color=True, app=ops.appearance(bias))
    fw0 = np_forward(net, feat, dir01, app=(w3e, emb, np.zeros(n, np.int64)))
    assert norm_rel(rgb_r0.float().cpu().numpy().reshape(3, n).T, fw0["rgb"]) < tol

    # backward
    ls = 1.0 if prec == "fp32" else 1024.0
    G_rgb = rng.normal(scale=1e-3, size=(n, 3)) * ls
    if prec == "fp16":
        G_rgb = G_rgb.astype(np.float16).astype(np.float64)
    dL_ds = rng.normal(scale=1e-3, size=n).astype(np.float32)
    t = rng.uniform(0.2, 2.0, size=n).astype(np.float32)
    grad = torch.zeros(net.size, dtype=torch.float32, device="cuda")
    dz = torch.full((64 * n,), float("nan"), dtype=torch.float32, device="cuda")
    dfeat = ops.mlp_backward(params, n_grid, n, feat_d, n, dir_d, n, torch.from_numpy(t).cuda(),
                             torch.from_numpy(np.ascontiguousarray(G_rgb.T.ravel())).to(dtype).cuda(),
                             torch.from_numpy(dL_ds).cuda(), ls, grad, app=ops.appearance(bias, img_d, dz))
    g_w3e, g_emb = ops.appearance_backward(n, n, img_d, dz, w3e_d, emb_d)
    r16 = prec == "fp16"
    fwr = np_forward(net, feat, dir01, round16=r16, app=app)
    dX, g_ref, gw_ref, ge_ref = np_backward(net, fwr, G_rgb, dL_ds, t, ls)
    got_w, got_e = g_w3e.cpu().numpy().reshape(64, 16), g_emb.cpu().numpy().reshape(n_img, 16)
    e_w, e_e = norm_rel(got_w, gw_ref), norm_rel(got_e, ge_ref)
    e_g, e_x = norm_rel(grad.cpu().numpy(), g_ref), norm_rel(dfeat.cpu().numpy().reshape(32, n).T, dX)
    print(f"[appearance bwd {prec}] W3e {e_w:.2e} emb {e_e:.2e} weights {e_g:.2e} dX {e_x:.2e}", flush=True)
    if prec == "fp32":
        assert max(e_w, e_e, e_g, e_x) < 1e-5
    else:
        # dz is formed from fp16 operands (the colour deltas) in fp32 and
        # summed per image in fp32: its error is that of the fp16 backward
        assert e_w < 1e-3 and e_e < 1e-3 and e_g < 1e-2 and e_x < 1e-2
    assert not got_e[2].any()  # no samples of image 2


def _trainer(tn, synthetic, dims, seed=51):
    blobs, cams, images = synthetic.blob_dataset(seed=seed, n_cams=10, width=64, height=64)
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
    cfg = tn.TrainingConfig()
    cfg.network.appearance_embedding_dims = dims
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 16)
    tr.load_images()
    return mgr, nerf, tr, cams


def _render(tn, synthetic, nerf, cam):
    buf = tn.CPURenderBuffer()
    tn.Renderer().submit(tn.RenderRequest(synthetic.to_pyturbonerf_cameras(tn, [cam])[0], [tn.Renderable(nerf)], buf))
    return np.array(buf.get_rgba())


def test_trainer_with_appearance_embedding(cuda, tmp_path):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    mgr, nerf, tr, cams = _trainer(tn, synthetic, 16)
    assert nerf.n_appearances == 10
    p0 = tr.get_params()
    n_grid = len(p0) - 9408 - 1024 - 160
    emb0 = p0[-160:].reshape(10, 16)
    assert (emb0 >= 0).all() and (emb0 < 1).all() and emb0.std() > 0.2  # uniform [0, 1)
    first = tr.train_step().loss

    # renders take image 0's embedding (checked while the colours are far
    # from the sigmoid's saturation, right after the first step)
    cam = cams[0]
    img_a = _render(tn, synthetic, nerf, cam)
    m = img_a[..., 3] > 0.05
    ratio = img_a[..., :3][m] / img_a[..., 3:4][m]
    assert m.any() and ratio.max() < 0.99
    p = tr.get_params()
    q = p.copy()
    q[-160 + 16:-160 + 32] += 0.5  # image 1
    tr.set_params(q)
    assert np.array_equal(_render(tn, synthetic, nerf, cam), img_a)
    q[-160:-160 + 16] += 0.5  # image 0
    tr.set_params(q)
    assert np.abs(_render(tn, synthetic, nerf, cam) - img_a).max() > 1e-3
    tr.set_params(p)

    for k in range(6):
        tr.train_steps(16, read_metrics=False, graph=k % 2 == 1)
        tr.update_occupancy_grid(nerf.training_step)
    last = tr.train_step().loss
    p1 = tr.get_params()
    moved = np.abs(p1[-160:] - p0[-160:]).reshape(10, 16).max(axis=1)
    print(f"[appearance] loss {first:.4f} -> {last:.4f}; embedding moves {np.round(moved, 4)}", flush=True)
    assert last < first and (moved > 0).all()
    w3e_moved = np.abs(p1[n_grid + 9408:n_grid + 9408 + 1024] - p0[n_grid + 9408:n_grid + 9408 + 1024]).max()
    assert w3e_moved > 0

    # snapshot with the embedding: the reference's header counts, exact reload
    img_b = _render(tn, synthetic, nerf, cam)
    path = str(tmp_path / "emb.turbo")
    tn.FileManager.save(nerf, path)
    head = open(path, "rb").read(300)
    n_color, n_density, n_emb = np.frombuffer(head[276:288], np.uint32)
    n_appearances = np.frombuffer(head[296:300], np.uint32)[0]
    assert (n_color, n_emb, n_appearances) == (8192, 160, 10) and n_density == n_grid + 3072
    mgr2 = tn.NeRFManager()
    other = mgr2.create()
    tn.FileManager.load(other, path)
    assert other.n_appearances == 10
    assert np.array_equal(_render(tn, synthetic, other, cam), img_b)
    tr.teardown()
