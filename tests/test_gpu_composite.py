"""GPU parity of compositing + smooth-L1 + backward (64 rays, geometric(48)
samples per ray clipped to [1, 1024], sigma ~ lognormal(0, 1.5), colours and
targets uniform). Checked against the oracle's fp64 restatement of the
reference kernel and against FP64 central finite differences (h = 1e-6).
Tolerances: loss rtol 1e-5; gradients norm-wise relative error <= 1e-4 for the
compensated-fp32 kernel, <= 1e-9 for its fp64 oracle mode, <= 2e-3 for fp16
inputs/outputs; finite differences <= 1e-3."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def make_batch(seed=7, n_rays=64, mean=48, dt_scale=4.0 / 48):
    rng = np.random.default_rng(seed)
    steps = np.clip(rng.geometric(1.0 / mean, n_rays), 1, 1024).astype(np.uint32)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    ns = int(steps.sum())
    sigma = rng.lognormal(0.0, 1.5, ns)
    s = (np.log(sigma) + 1.0).astype(np.float32)
    color = rng.uniform(0, 1, 3 * ns).astype(np.float32)
    dt = (rng.uniform(0.5, 1.5, ns) * dt_scale).astype(np.float32)
    bg = rng.uniform(0, 1, 3 * n_rays).astype(np.float32)
    target = rng.uniform(0, 1, 4 * n_rays).astype(np.float32)
    return dict(n_rays=n_rays, steps=steps, offs=offs, ns=ns, s=s, color=color, dt=dt, bg=bg, target=target)


def norm_rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def oracle_f64(b, s=None, color=None):
    """Per-ray arrays have stride n_rays, per-sample arrays stride ns."""
    import oracle

    nr, ns = b["n_rays"], b["ns"]
    s = b["s"] if s is None else s
    color = b["color"] if color is None else color
    loss, dLd, dLc = oracle.composite_train(nr, nr, ns, b["steps"], b["offs"], b["bg"].astype(np.float64),
                                            b["target"].astype(np.float64), s.astype(np.float64),
                                            color.astype(np.float64), b["dt"].astype(np.float64), 1.0, 1e-4, True)
    return loss.reshape(4, nr), dLd, dLc.reshape(3, ns)


def run_gpu(b, dtype, fp64_accumulate=False):
    import torch

    from turbonerf_amd import ops

    ns, nr = b["ns"], b["n_rays"]
    T = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    info = torch.tensor([nr, ns, 0, 0], dtype=torch.int32, device="cuda")
    loss, dLd, dLc = ops.composite_train(info, nr, nr, ns, T(b["steps"].view(np.int32), torch.int32),
                                         T(b["offs"].view(np.int32), torch.int32), T(b["bg"]), T(b["target"]),
                                         T(b["s"], dtype), T(b["color"], dtype), T(b["dt"]), 1.0, 1e-4,
                                         fp64_accumulate)
    torch.cuda.synchronize()
    return (loss.cpu().numpy().reshape(4, nr), dLd.cpu().numpy().astype(np.float64),
            dLc.float().cpu().numpy().reshape(3, ns).astype(np.float64))


def test_composite_fp32_vs_oracle(cuda):
    b = make_batch()
    print(f"[composite] rays={b['n_rays']} samples={b['ns']}", flush=True)
    l_ref, d_ref, c_ref = oracle_f64(b)
    l, d, c = run_gpu(b, __import__("torch").float32)
    np.testing.assert_allclose(l, l_ref, rtol=1e-5, atol=1e-7)
    ed, ec = norm_rel(d, d_ref), norm_rel(c, c_ref)
    print(f"[composite fp32] rel err dL/ds={ed:.2e} dL/dc={ec:.2e}", flush=True)
    assert ed < 1e-4 and ec < 1e-4


def test_composite_fp64_mode(cuda):
    b = make_batch(seed=11)
    l_ref, d_ref, c_ref = oracle_f64(b)
    l, d, c = run_gpu(b, __import__("torch").float32, fp64_accumulate=True)
    np.testing.assert_allclose(l, l_ref, rtol=1e-6, atol=1e-9)
    # inputs and outputs are fp32; the accumulation is fp64
    assert norm_rel(d, d_ref) < 1e-6 and norm_rel(c, c_ref) < 1e-6


def test_composite_fp16(cuda):
    import torch

    b = make_batch(seed=3)
    b16 = dict(b)
    b16["s"] = b["s"].astype(np.float16).astype(np.float32)
    b16["color"] = b["color"].astype(np.float16).astype(np.float32)
    l_ref, d_ref, c_ref = oracle_f64(b16)
    l, d, c = run_gpu(b, torch.float16)
    np.testing.assert_allclose(l, l_ref, rtol=1e-4, atol=1e-6)
    assert norm_rel(d, d_ref) < 1e-4 and norm_rel(c, c_ref) < 2e-3


def test_composite_fd(cuda):
    """FP64 central differences of the oracle's fp64 forward loss."""
    import torch

    b = make_batch(seed=5)
    nr = b["n_rays"]
    l, d, c = run_gpu(b, torch.float32)

    def total(s, color):
        loss, _, _ = oracle_f64(b, s, color)
        return loss[:3].sum()  # rgb terms carry the gradient (alpha is reported only)

    rng = np.random.default_rng(0)
    idx = rng.choice(b["ns"], 80, replace=False)
    h = 1e-6
    s64 = b["s"].astype(np.float64)
    c64 = b["color"].astype(np.float64)
    fd_s, an_s, fd_c, an_c = [], [], [], []
    for i in idx:
        sp, sm = s64.copy(), s64.copy()
        sp[i] += h; sm[i] -= h
        fd_s.append((total(sp, c64) - total(sm, c64)) / (2 * h) / nr)
        an_s.append(d[i])
        cp, cm = c64.copy(), c64.copy()
        cp[i] += h; cm[i] -= h
        fd_c.append((total(s64, cp) - total(s64, cm)) / (2 * h) / nr)
        an_c.append(c[0, i])
    # total() sums per-ray losses; the kernel's gradient is of their mean over rays
    fd_s, an_s, fd_c, an_c = map(np.array, (fd_s, an_s, fd_c, an_c))
    es, ec = norm_rel(an_s, fd_s), norm_rel(an_c, fd_c)
    print(f"[composite fd] rel err dL/ds={es:.2e} dL/dc={ec:.2e}", flush=True)
    assert es < 1e-3 and ec < 1e-3


@pytest.mark.parametrize("dt_scale", [0.5 / 400, 0.015])
def test_composite_long_rays(cuda, dt_scale):
    """Rays of about 400 samples (up to 1024), each spanning many 64-sample
    chunks: with little optical depth (no ray stops) and with transmittance
    falling below 1e-4 a few hundred samples in, inside a chunk the ray
    entered from an earlier one."""
    import torch

    b = make_batch(seed=21, n_rays=24, mean=400, dt_scale=dt_scale)
    l_ref, d_ref, c_ref = oracle_f64(b)
    l, d, c = run_gpu(b, torch.float32)
    np.testing.assert_allclose(l, l_ref, rtol=1e-5, atol=1e-7)
    ed, ec = norm_rel(d, d_ref), norm_rel(c, c_ref)
    print(f"[composite long dt={dt_scale:.4g}] rel err dL/ds={ed:.2e} dL/dc={ec:.2e}", flush=True)
    assert ed < 1e-4 and ec < 1e-4
    l, d, c = run_gpu(b, torch.float32, fp64_accumulate=True)
    np.testing.assert_allclose(l, l_ref, rtol=1e-6, atol=1e-9)
    assert norm_rel(d, d_ref) < 1e-6 and norm_rel(c, c_ref) < 1e-6


def test_step_metrics(cuda):
    """Mean loss over info[0] rays (entries past it ignored), mean distortion
    loss or 0, and the device sample counter."""
    import torch

    from turbonerf_amd._capi import call, ptr, stream_ptr

    rng = np.random.default_rng(4)
    stride, n = 3000, 2345
    loss = rng.uniform(0, 1, 4 * stride).astype(np.float32)
    dist = rng.uniform(0, 1, stride).astype(np.float32)
    info = torch.tensor([n, 123456, 0, 0], dtype=torch.int32, device="cuda")
    total = torch.tensor([1000], dtype=torch.int64, device="cuda")
    out = torch.full((2,), 7.0, dtype=torch.float32, device="cuda")
    ld, dd = torch.from_numpy(loss).cuda(), torch.from_numpy(dist).cuda()
    call("tn_step_metrics", ptr(info), stride, ptr(ld), ptr(dd), ptr(out), ptr(total), None, stream_ptr())
    o = out.cpu().numpy()
    ref = loss.astype(np.float64).reshape(4, stride)[:, :n].sum() / n
    assert abs(o[0] - ref) <= 1e-6 * ref and abs(o[1] - dist[:n].astype(np.float64).mean()) <= 1e-6
    assert int(total.cpu()[0]) == 1000 + 123456
    call("tn_step_metrics", ptr(info), stride, ptr(ld), None, ptr(out), None, None, stream_ptr())
    o = out.cpu().numpy()
    assert o[1] == 0.0 and abs(o[0] - ref) <= 1e-6 * ref and int(total.cpu()[0]) == 1000 + 123456
