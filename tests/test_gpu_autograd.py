"""The adjoint kernels behind torch.autograd.Function (turbonerf_amd.autograd):

* hash grid: the forward is linear in the table, so <F v, w> = <v, F^T w>
  exactly up to rounding (dot-product test, 1e-5), and autograd's table
  gradient equals the scatter-add kernel's (1e-6: atomics reorder sums);
* compositing and distortion: autograd gradients (scaled by an upstream 2.5)
  against the fp64 oracle (norm-wise 1e-5 / 1e-4);
* MLP: autograd's vector-Jacobian product against the fp64 numpy backward
  (1e-5, fp32);
* the whole chain table -> hash grid -> MLP -> compositing loss, fp32 on the
  GPU through autograd, against the same chain in fp64 through the oracle
  (loss rtol 1e-5, d loss / d table and d loss / d weights norm-wise 1e-4).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def norm_rel(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _rays(seed, n_rays=64, mean=48):
    rng = np.random.default_rng(seed)
    steps = np.clip(rng.geometric(1 / mean, n_rays), 1, 1024).astype(np.uint32)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    S = int(steps.sum())
    return rng, steps, offs, S


def test_hashgrid_adjoint_and_autograd(cuda):
    import torch

    from turbonerf_amd import ops
    from turbonerf_amd.autograd import HashGridEncoding

    g = ops.hashgrid_layout()
    rng = np.random.default_rng(31)
    n = 3000
    pos = torch.from_numpy(rng.uniform(0, 1, (3, n)).astype(np.float32)).cuda()
    v = torch.from_numpy(rng.normal(size=g.n_params).astype(np.float32)).cuda()
    w = torch.from_numpy(rng.normal(size=(32, n)).astype(np.float32)).cuda()
    Fv = HashGridEncoding.apply(v, pos, g)
    table = v.clone().requires_grad_(True)
    out = HashGridEncoding.apply(table, pos, g)
    (out * w).sum().backward()
    lhs = float((Fv.double() * w.double()).sum())
    rhs = float((v.double() * table.grad.double()).sum())
    print(f"[autograd hashgrid] <Fv,w> {lhs:.6e} <v,F^T w> {rhs:.6e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
    direct = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, pos.contiguous(), n, w.contiguous(), direct)
    # same kernel, atomics in a different order
    assert norm_rel(table.grad.cpu().numpy(), direct.cpu().numpy()) < 1e-6


def test_composite_and_distortion_autograd(cuda):
    import torch

    import oracle
    from turbonerf_amd.autograd import CompositeLoss, DistortionLoss

    rng, steps, offs, S = _rays(32)
    nr = len(steps)
    s = rng.normal(1.0, 1.5, S).astype(np.float32)
    c = rng.uniform(0, 1, (3, S)).astype(np.float32)
    dt = rng.uniform(2e-3, 2e-2, S).astype(np.float32)
    bg = rng.uniform(0, 1, (3, nr)).astype(np.float32)
    tgt = rng.uniform(0, 1, (4, nr)).astype(np.float32)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    st, of = T(steps.view(np.int32)), T(offs.view(np.int32))
    dens = T(s).requires_grad_(True)
    col = T(c).requires_grad_(True)
    loss = CompositeLoss.apply(dens, col, T(dt), st, of, T(bg), T(tgt))
    (2.5 * loss).backward()
    lr, dLd, dLc = oracle.composite_train(nr, nr, S, steps, offs, bg.ravel(), tgt.ravel(), s, c.ravel(), dt)
    ref_loss = lr.reshape(4, nr)[:3].sum() / nr
    e1, e2 = norm_rel(dens.grad.cpu().numpy(), 2.5 * dLd), norm_rel(col.grad.cpu().numpy(), 2.5 * dLc)
    print(f"[autograd composite] loss {float(loss):.6e} vs {ref_loss:.6e}; grads {e1:.2e} {e2:.2e}", flush=True)
    assert abs(float(loss) - ref_loss) <= 1e-5 * abs(ref_loss)
    assert e1 < 1e-5 and e2 < 1e-5
    # distortion over the same rays with normalised midpoints and steps
    m = np.zeros(S, np.float32)
    d = np.zeros(S, np.float32)
    for i in range(nr):
        k, o = int(steps[i]), int(offs[i])
        w_ = rng.uniform(0.5, 1.5, k)
        t0 = np.concatenate([[0], np.cumsum(w_)[:-1]])
        span = t0[-1] + w_[-1]
        m[o:o + k] = (t0 + 0.5 * w_) / span
        d[o:o + k] = w_ / span
    dens2 = T(s).requires_grad_(True)
    dl = DistortionLoss.apply(dens2, T(m), T(d), st, of, 1e-3)
    (2.5 * dl).backward()
    l_ref, g_ref = oracle.distortion(nr, steps, offs, s.astype(np.float64), m.astype(np.float64),
                                     d.astype(np.float64), 1e-3)
    e3 = norm_rel(dens2.grad.cpu().numpy(), 2.5 * g_ref)
    print(f"[autograd distortion] loss {float(dl):.6e} vs {l_ref.sum():.6e}; grad {e3:.2e}", flush=True)
    assert abs(float(dl) - l_ref.sum()) <= 1e-5 * abs(l_ref.sum())
    assert e3 < 1e-4


def test_mlp_autograd(cuda):
    import torch

    from nets import backward as np_backward
    from nets import forward as np_forward
    from nets import random_net
    from turbonerf_amd.autograd import NerfMLP

    rng = np.random.default_rng(33)
    n = 1000
    net = random_net(33)
    feat = rng.normal(scale=0.5, size=(n, 32))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dir01 = (0.5 * d + 0.5).astype(np.float32)
    a = rng.normal(size=n)
    b = rng.normal(size=(n, 3))
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    W = T(net).requires_grad_(True)
    X = T(feat.T).requires_grad_(True)
    dens, rgb = NerfMLP.apply(W, X, T(dir01.T))
    ((dens * T(a)).sum() + (rgb * T(b.T)).sum()).backward()
    fw = np_forward(net, feat, dir01)
    dX, g = np_backward(net, fw, b, a, np.ones(n), 1.0)
    e1, e2 = norm_rel(W.grad.cpu().numpy(), g), norm_rel(X.grad.cpu().numpy().T, dX)
    print(f"[autograd mlp] dW {e1:.2e} dX {e2:.2e}", flush=True)
    assert e1 < 1e-5 and e2 < 1e-5


def test_chain_table_to_loss_vs_fp64(cuda):
    import torch

    import oracle
    from turbonerf_amd import ops
    from turbonerf_amd.autograd import CompositeLoss, HashGridEncoding, NerfMLP
    from nets import random_net

    g = ops.hashgrid_layout()
    rng, steps, offs, S = _rays(34, n_rays=48, mean=40)
    nr = len(steps)
    # samples along random rays through the unit cube
    pos = np.zeros((3, S), np.float32)
    dirs = np.zeros((3, S), np.float32)
    for i in range(nr):
        k, o = int(steps[i]), int(offs[i])
        p0 = rng.uniform(0.2, 0.8, 3)
        dv = rng.normal(size=3)
        dv /= np.linalg.norm(dv)
        tt = np.arange(k) * 0.004
        pos[:, o:o + k] = np.clip(p0[:, None] + dv[:, None] * tt, 0, 1)
        dirs[:, o:o + k] = (0.5 * dv + 0.5)[:, None]
    dt = np.full(S, 0.004, np.float32)
    table = rng.uniform(-1.0, 1.0, g.n_params).astype(np.float32)
    net = random_net(35).astype(np.float32)
    bg = rng.uniform(0, 1, (3, nr)).astype(np.float32)
    tgt = rng.uniform(0, 1, (4, nr)).astype(np.float32)
    T = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    tab = T(table).requires_grad_(True)
    W = T(net).requires_grad_(True)
    feat = HashGridEncoding.apply(tab, T(pos), g)
    dens, rgb = NerfMLP.apply(W, feat, T(dirs))
    loss = CompositeLoss.apply(dens, rgb, T(dt), T(steps.view(np.int32)), T(offs.view(np.int32)), T(bg), T(tgt))
    loss.backward()
    # fp64 chain through the oracle
    L = oracle.lib()
    f64 = np.zeros(32 * S)
    L.oc_hashgrid_fwd_f64(ctypes.addressof(g), S, S, np.ascontiguousarray(pos.ravel()), table.astype(np.float64), f64)
    f64 = f64.reshape(32, S)
    d_ref, rgb_ref, _, _ = oracle.mlp_train(net.astype(np.float64), f64, dirs, None, None, None, 1.0)
    lr, dLd, dLc = oracle.composite_train(nr, nr, S, steps, offs, bg.ravel(), tgt.ravel(), d_ref, rgb_ref.ravel(),
                                          dt)
    ref_loss = lr.reshape(4, nr)[:3].sum() / nr
    _, _, dfeat, dW = oracle.mlp_train(net.astype(np.float64), f64, dirs, dLc.reshape(3, S), dLd.astype(np.float32),
                                       np.ones(S, np.float32), 1.0)
    dtab = np.zeros(g.n_params)
    L.oc_hashgrid_bwd_f64(ctypes.addressof(g), S, S, np.ascontiguousarray(pos.ravel()), dfeat.ravel(), dtab)
    e_t, e_w = norm_rel(tab.grad.cpu().numpy(), dtab), norm_rel(W.grad.cpu().numpy(), dW)
    print(f"[autograd chain] loss {float(loss):.6e} vs {ref_loss:.6e}; dtable {e_t:.2e} dW {e_w:.2e}", flush=True)
    assert abs(float(loss) - ref_loss) <= 1e-5 * abs(ref_loss)
    assert e_t < 1e-4 and e_w < 1e-4
