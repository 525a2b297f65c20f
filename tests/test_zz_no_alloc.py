"""The allocation-free compositing entry point (tn_composite_train_noalloc:
the compositing kernels with their scratch in a static per-GPU device array)
against tn_composite_train with a caller workspace, through the C ABI: two
calls in a row, the second on a smaller batch, each bit-exact against the
workspace entry point on the same inputs, with fp32 and fp16 network outputs
and the ray-count gradient scaling on; a batch whose scratch exceeds the
static array is refused with TN_ERR_OUT_OF_MEMORY before any GPU work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(seed, n_rays, mean):
    rng = np.random.default_rng(seed)
    steps = np.clip(rng.geometric(1.0 / mean, n_rays), 1, 1024).astype(np.uint32)
    offs = np.concatenate([[0], np.cumsum(steps)[:-1]]).astype(np.uint32)
    ns = int(steps.sum())
    s = (np.log(rng.lognormal(0.0, 1.5, ns)) + 1.0).astype(np.float32)
    return dict(n_rays=n_rays, ns=ns, steps=steps, offs=offs, s=s,
                color=rng.uniform(0, 1, 3 * ns).astype(np.float32),
                dt=(rng.uniform(0.5, 1.5, ns) / 12.0).astype(np.float32),
                bg=rng.uniform(0, 1, 3 * n_rays).astype(np.float32),
                target=rng.uniform(0, 1, 4 * n_rays).astype(np.float32))


def _run(b, dtype, noalloc):
    import torch

    from turbonerf_amd import _capi, ops
    from turbonerf_amd._capi import call, ptr, stream_ptr

    nr, ns = b["n_rays"], b["ns"]
    T = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).to(dt).cuda()
    info = torch.tensor([nr, ns, 0, 0], dtype=torch.int32, device="cuda")
    code = 1 if dtype == torch.float16 else 0
    steps, offs = T(b["steps"].view(np.int32), torch.int32), T(b["offs"].view(np.int32), torch.int32)
    bg, target, s, color, dt = T(b["bg"]), T(b["target"]), T(b["s"], dtype), T(b["color"], dtype), T(b["dt"])
    loss = torch.full((4 * nr,), 7.0, dtype=torch.float32, device="cuda")
    dLd = torch.full((ns,), 7.0, dtype=torch.float32, device="cuda")
    dLc = torch.full((3 * ns,), 7.0, dtype=dtype, device="cuda")
    args = (code, ptr(info), nr, nr, ns, 1024.0, 1, 1e-4, ptr(steps), ptr(offs), ptr(bg), ptr(target), ptr(s),
            ptr(color), ptr(dt), ptr(loss), ptr(dLd), ptr(dLc))
    if noalloc:
        call("tn_composite_train_noalloc", *args, stream_ptr())
    else:
        nb = ops.query("tn_composite_scratch_bytes", nr, ns)
        scratch = torch.empty(nb // 4 + 64, dtype=torch.float32, device="cuda")
        call("tn_composite_train", *args, ptr(scratch), nb, stream_ptr())
    torch.cuda.synchronize()
    return loss.cpu().numpy(), dLd.cpu().numpy(), dLc.float().cpu().numpy()


@pytest.mark.parametrize("half", [False, True])
def test_noalloc_composite_matches_workspace_entry_point(cuda, half):
    import torch

    dtype = torch.float16 if half else torch.float32
    for k, (n_rays, mean) in enumerate(((3000, 90), (700, 40))):
        b = _batch(31 + k, n_rays, mean)
        print(f"[noalloc] call {k}: rays {n_rays} samples {b['ns']} fp16={half}", flush=True)
        ref = _run(b, dtype, False)
        got = _run(b, dtype, True)
        for r, g in zip(ref, got):
            assert np.array_equal(r.view(np.uint32), g.view(np.uint32))
        assert np.abs(ref[1]).max() > 0


def test_noalloc_composite_rejects_oversized_batches(cuda):
    from turbonerf_amd import _capi, ops

    limit = ops.query("tn_composite_train_noalloc_max_bytes")
    assert limit == 32 << 20
    n_rays = 1 << 20  # 56 MB of per-ray scratch
    assert ops.query("tn_composite_scratch_bytes", n_rays, 1 << 20) > limit
    rc = _capi.lib().tn_composite_train_noalloc(0, None, n_rays, n_rays, 1 << 20, 1.0, 0, 1e-4, None, None, None,
                                                None, None, None, None, None, None, None, None)
    assert rc == 4  # TN_ERR_OUT_OF_MEMORY ("workspace too small")
