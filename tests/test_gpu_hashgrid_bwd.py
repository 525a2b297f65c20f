"""The binned hash-grid backward (counting sort into 4096-entry bins, 64-bit
fixed-point sums in LDS) against the oracle's fp64 scatter-add and the atomic
kernel, on the default 16-level layout (T = 2^19):

* samples along rays (runs of equal coarse entries merged before sorting),
  fp32 and fp16 feature gradients, a share of samples with zero gradient
  (their records carry zeros: the record counts follow from the positions
  alone), batch size from the device count (info);
* the same in two calls, the plan from the positions first
  (tn_hashgrid_backward_plan) and then the gradient passes
  (tn_hashgrid_backward_planned), as the trainer runs them, or the records
  of two sample ranges scattered apart and summed once: equal to the
  one-call backward bit for bit;
* random points (no merging: level 0's single bin holds 16 chunks of
  records, summed in parts);
* accumulation into a non-zero table, values spanning 1e-6..1e3 (the
  fixed-point scale follows each bin's largest value), and a non-finite
  gradient (the bin falls back to float sums and the inf reaches the table);
* FP64 central differences (h = 1e-6) of the oracle's fp64 forward on
  samples along 64 rays (geometric(48) samples per ray): the binned
  gradient of the loss sum(dfeat * features) at table entries those samples
  touch;
* Adam on the table fused into the sum passes
  (tn_hashgrid_backward_planned_adam): the parameters, moments, steps and
  fp16 copy of the unfused backward followed by tn_adam_step, bit for bit,
  with split bins and a non-finite gradient.

Bar: norm-wise 1e-6 against fp64 (atomic kernel), 2e-6 (binned); for the
binned kernel also per entry |err| <= 6e-6 sum|w g| (fp32 products and run
sums, 2e-6, and each record's value rounded to 17 significant bits, 2^-18)
plus 6.4e-8 max|g| of the level (the fixed-point rounding: 2^-S per record,
at most 2^-31 of a bin's largest record, which is a run of at most 64
samples)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def norm_rel(a, b):
    a, b = np.ravel(np.asarray(a, np.float64)), np.ravel(np.asarray(b, np.float64))
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _rays(rng, n_rays, per, spacing):
    o = rng.normal(size=(n_rays, 3))
    o = 0.5 + 1.2 * o / np.linalg.norm(o, axis=1, keepdims=True)
    tgt = rng.uniform(0.3, 0.7, size=(n_rays, 3))
    d = tgt - o
    ln = np.linalg.norm(d, axis=1, keepdims=True)
    d /= ln
    t = (ln - 0.45) + spacing * np.arange(per)[None, :]
    p = o[:, None, :] + t[..., None] * d[:, None, :]
    return np.clip(p.reshape(-1, 3), 0.0, 1.0).astype(np.float32).T.copy()  # [3][n]


def _oracle_bwd(g, pos, dfeat):
    import oracle

    n = pos.shape[1]
    out = np.zeros(g.n_params)
    oracle.lib().oc_hashgrid_bwd_f64(ctypes.addressof(g), n, n, np.ascontiguousarray(pos.ravel()),
                                     np.ascontiguousarray(dfeat, np.float64).ravel(), out)
    return out


def _level_of_entry(g):
    lv = np.zeros(g.n_params // 2, np.int64)
    for l in range(g.n_levels):
        lv[g.level_offset[l]:g.level_offset[l] + g.level_size[l]] = l
    return np.repeat(lv, 2)


def _check(got, ref, tag, g=None, ref_abs=None, dfeat=None, binned=True):
    rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)
    msg = f"[hash bwd {tag}] rel {rel:.2e}"
    assert rel < (2e-6 if binned else 1e-6), msg
    if ref_abs is not None:
        gmax = np.array([np.abs(dfeat[2 * l:2 * l + 2]).max() for l in range(g.n_levels)])
        bar = 6e-6 * ref_abs + 6.4e-8 * gmax[_level_of_entry(g)]
        worst = float(np.max(np.abs(got - ref) / np.maximum(bar, 1e-300)))
        msg += f" worst/bar {worst:.3f}"
        assert worst <= 1.0, msg
    print(msg, flush=True)


@pytest.mark.parametrize("half", [False, True])
def test_binned_rays_vs_oracle(cuda, half):
    import torch

    from turbonerf_amd import ops

    rng = np.random.default_rng(11)
    g = ops.hashgrid_layout()
    n_rays, per = 600, 160
    pos = _rays(rng, n_rays, per, 0.0017)
    n = pos.shape[1]
    dfeat = rng.normal(0.0, 0.01, size=(32, n)).astype(np.float16 if half else np.float32)
    dfeat[:, (np.arange(n) % per) >= 100] = 0  # samples behind the surface carry no gradient
    dfeat_used = dfeat.astype(np.float64)
    ref = _oracle_bwd(g, pos, dfeat_used)
    ref_abs = _oracle_bwd(g, pos, np.abs(dfeat_used))
    dt = torch.float16 if half else torch.float32
    B = n + 77  # capacity above the batch; the count comes from info
    pos_d = torch.zeros(3, B, dtype=torch.float32, device="cuda")
    pos_d[:, :n] = torch.from_numpy(pos).cuda()
    df_d = torch.zeros(32, B, dtype=dt, device="cuda")
    df_d[:, :n] = torch.from_numpy(dfeat).cuda()
    info = torch.tensor([0, n, 0, 0], dtype=torch.int32, device="cuda")
    got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, B, pos_d.view(-1), B, df_d.view(-1), got, info=info, binned=True)
    atom = torch.zeros_like(got)
    ops.hashgrid_backward(g, B, pos_d.view(-1), B, df_d.view(-1), atom, info=info)
    got_h, atom_h = got.cpu().numpy().astype(np.float64), atom.cpu().numpy().astype(np.float64)
    _check(got_h, ref, f"rays {'fp16' if half else 'fp32'}", g, ref_abs, dfeat_used)
    _check(atom_h, ref, "rays atomic", binned=False)
    # untouched entries stay exactly zero in both
    assert np.all(got_h[ref == 0] == 0)


def test_binned_random_points_chunked_bins(cuda):
    import torch

    from turbonerf_amd import ops

    rng = np.random.default_rng(12)
    g = ops.hashgrid_layout()
    n = 65536  # level 0 (4096 entries, one bin) receives 8 records per sample: 16 chunks
    pos = rng.uniform(0.0, 1.0, size=(3, n)).astype(np.float32)
    dfeat = rng.normal(0.0, 1.0, size=(32, n)).astype(np.float32)
    ref = _oracle_bwd(g, pos, dfeat)
    got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, torch.from_numpy(pos).cuda().view(-1), n, torch.from_numpy(dfeat).cuda().view(-1),
                          got, binned=True)
    _check(got.cpu().numpy().astype(np.float64), ref, "random", g, _oracle_bwd(g, pos, np.abs(dfeat)),
           dfeat.astype(np.float64))


def test_binned_accumulates_wide_range_and_nonfinite(cuda):
    import torch

    from turbonerf_amd import ops

    rng = np.random.default_rng(13)
    g = ops.hashgrid_layout()
    pos = _rays(rng, 200, 100, 0.003)
    n = pos.shape[1]
    # magnitudes from 1e-6 to 1e3 across samples
    dfeat = (rng.normal(size=(32, n)) * 10.0 ** rng.uniform(-6, 3, size=(1, n))).astype(np.float32)
    ref = _oracle_bwd(g, pos, dfeat)
    pos_d = torch.from_numpy(pos).cuda().view(-1)
    df_d = torch.from_numpy(dfeat).cuda().view(-1)
    got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, pos_d, n, df_d, got, binned=True)
    _check(got.cpu().numpy().astype(np.float64), ref, "wide range", g, _oracle_bwd(g, pos, np.abs(dfeat)),
           dfeat.astype(np.float64))
    # accumulation into a non-zero table: table + increment
    base = rng.normal(0.0, 1.0, size=g.n_params).astype(np.float32)
    acc = torch.from_numpy(base.copy()).cuda()
    ops.hashgrid_backward(g, n, pos_d, n, df_d, acc, binned=True)
    _check(acc.cpu().numpy().astype(np.float64) - base, ref, "accumulate")
    assert np.all(acc.cpu().numpy()[ref == 0] == base[ref == 0])
    # one infinite gradient: its entries become inf, the rest stay finite
    dinf = dfeat.copy()
    dinf[0, 5] = np.inf
    got2 = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, pos_d, n, torch.from_numpy(dinf).cuda().view(-1), got2, binned=True)
    g2 = got2.cpu().numpy()
    ref2 = _oracle_bwd(g, pos, dinf.astype(np.float64))
    assert np.array_equal(np.isinf(g2), np.isinf(ref2)) and np.isinf(g2).sum() > 0
    fin = np.isfinite(ref2)
    _check(np.where(fin, g2, 0.0).astype(np.float64), np.where(fin, ref2, 0.0), "with inf")


@pytest.mark.parametrize("log2_T", [15, 21])
def test_binned_table_sizes(cuda, log2_T):
    """Tables of 2^15 (most levels dense-sized) and 2^21 entries per level
    (11 hashed levels of 512 bins: 5,717 bins in all)."""
    import torch

    from turbonerf_amd import _capi, ops

    rng = np.random.default_rng(14 + log2_T)
    g = ops.hashgrid_layout(log2_T=log2_T)
    pos = _rays(rng, 300, 120, 0.002)
    n = pos.shape[1]
    assert _capi.query("tn_hashgrid_backward_workspace_bytes", n, g) > 0
    dfeat = rng.normal(0.0, 0.05, size=(32, n)).astype(np.float32)
    ref = _oracle_bwd(g, pos, dfeat)
    got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, torch.from_numpy(pos).cuda().view(-1), n, torch.from_numpy(dfeat).cuda().view(-1),
                          got, binned=True)
    _check(got.cpu().numpy().astype(np.float64), ref, f"T=2^{log2_T}", g, _oracle_bwd(g, pos, np.abs(dfeat)),
           dfeat.astype(np.float64))


def test_binned_is_reproducible(cuda):
    """Two runs over random points (level 0's bin split over 16 blocks) give
    the same gradient bit for bit."""
    import torch

    from turbonerf_amd import ops

    rng = np.random.default_rng(16)
    g = ops.hashgrid_layout()
    n = 65536
    pos = torch.from_numpy(rng.uniform(0, 1, size=(3, n)).astype(np.float32)).cuda().view(-1)
    df = torch.from_numpy(rng.normal(0.0, 1.0, size=(32, n)).astype(np.float32)).cuda().view(-1)
    runs = []
    for _ in range(2):
        got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
        ops.hashgrid_backward(g, n, pos, n, df, got, binned=True)
        runs.append(got.cpu().numpy())
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))


@pytest.mark.parametrize("groups,level_blocks", [(2, 128), (3, 40), (16, 8), (1, 256)])
def test_binned_level_groups_and_blocks_give_the_same_bits(cuda, groups, level_blocks):
    """Scattering and summing the levels in groups, and fewer count/scatter
    blocks per level, give the default launch's gradient bit for bit (rays
    with merged runs, plus random points whose level-0 bin is split)."""
    import torch

    from turbonerf_amd import _capi, ops

    lib = _capi.lib()
    rng = np.random.default_rng(17)
    g = ops.hashgrid_layout()
    pos = np.concatenate([_rays(rng, 512, 64, 0.004), rng.uniform(0, 1, size=(3, 32768)).astype(np.float32)], axis=1)
    n = pos.shape[1]
    pos_d = torch.from_numpy(np.ascontiguousarray(pos)).cuda().view(-1)
    df = torch.from_numpy(rng.normal(0.0, 1.0, size=(32, n)).astype(np.float32)).cuda().view(-1)
    runs = []
    try:
        for knobs in ({6: 1, 1: 128}, {6: groups, 1: level_blocks}):
            for k, v in knobs.items():
                assert lib.tn_tune_set(k, v) == 0
            got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
            ops.hashgrid_backward(g, n, pos_d, n, df, got, binned=True)
            runs.append(got.cpu().numpy())
    finally:
        assert lib.tn_tune_set(6, 1) == 0 and lib.tn_tune_set(1, 128) == 0
    assert np.abs(runs[0]).max() > 0
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))


@pytest.mark.parametrize("half", [False, True])
def test_planned_equals_one_call(cuda, half):
    """Plan from the positions, then the gradient passes: the one-call binned
    backward's gradient bit for bit (rays with merged runs, random points, a
    share of zero gradients, the batch bounded by a device count)."""
    import torch

    from turbonerf_amd import ops

    rng = np.random.default_rng(18)
    g = ops.hashgrid_layout()
    pos = np.concatenate([_rays(rng, 512, 64, 0.004), rng.uniform(0, 1, size=(3, 16384)).astype(np.float32)], axis=1)
    n = pos.shape[1]
    df = rng.normal(0.0, 1.0, size=(32, n)).astype(np.float32)
    df[:, rng.uniform(size=n) < 0.3] = 0.0
    pos_d = torch.from_numpy(np.ascontiguousarray(pos)).cuda().view(-1)
    df_d = torch.from_numpy(df).cuda().view(-1)
    if half:
        df_d = df_d.half()
    info = torch.tensor([0, n - 1000, 0, 0], dtype=torch.int32, device="cuda")
    # records scattered in two parts (tn_hashgrid_backward_scatter, the later
    # samples first) and summed once (tn_hashgrid_backward_accumulate): at
    # the middle count block, and at the last block, which lies past the
    # device count (its part scatters nothing but its zero bin maxima)
    part = ops.query("tn_hashgrid_backward_part_samples", n, g)
    blocks = -(-n // part)
    assert part % 256 == 0 and blocks >= 2
    splits = [None, None, part * (blocks // 2), part * (blocks - 1)]
    runs = []
    for k, split in enumerate(splits):
        got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
        ops.hashgrid_backward(g, n, pos_d, n, df_d, got, info=info, binned=k == 0, planned=k > 0, split=split)
        runs.append(got.cpu().numpy())
    print(f"[hash bwd planned half={half}] touched {np.count_nonzero(runs[0])}, parts of {part} samples", flush=True)
    assert np.abs(runs[0]).max() > 0
    for r in runs[1:]:
        assert np.array_equal(runs[0].view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("lp", [0, 1])
@pytest.mark.parametrize("nonfinite", [False, True])
def test_planned_adam_equals_backward_then_adam(cuda, lp, nonfinite):
    """Adam fused into the sum passes against tn_hashgrid_backward_planned
    then tn_adam_step over the table: random points (level 0's single bin
    holds 8 records per sample, summed in parts and finished by the last
    pass), rays (merged runs), a share of zero gradients, the batch bounded
    by a device count; with `nonfinite` an inf gradient puts level 0's split
    bin on float sums through the table-gradient buffer. Weights, moments,
    per-parameter steps and the fp16 copy agree bit for bit (that float-summed
    bin to rounding: its atomic sums take a different order in each run),
    and the gradient buffer is all zero afterwards in both."""
    import torch

    from turbonerf_amd import _capi, ops

    rng = np.random.default_rng(31)
    g = ops.hashgrid_layout()
    P = g.n_params
    pos = np.concatenate([_rays(rng, 256, 64, 0.004), rng.uniform(0, 1, size=(3, 16384)).astype(np.float32)], axis=1)
    n = pos.shape[1]
    df = rng.normal(0.0, 1.0, size=(32, n)).astype(np.float32) * 1e3
    df[:, rng.uniform(size=n) < 0.3] = 0.0
    if nonfinite:
        df[0, 20000] = np.inf
    pos_d = torch.from_numpy(np.ascontiguousarray(pos)).cuda().view(-1)
    df_d = torch.from_numpy(df).cuda().view(-1)
    info = torch.tensor([1000, n - 500, 0, 0], dtype=torch.int32, device="cuda")
    w0 = torch.from_numpy(rng.uniform(-1e-2, 1e-2, size=P).astype(np.float32)).cuda()
    m0 = torch.from_numpy(rng.normal(0.0, 1e-3, size=P).astype(np.float32)).cuda()
    v0 = torch.from_numpy(np.abs(rng.normal(0.0, 1e-6, size=P)).astype(np.float32)).cuda()
    s0 = torch.from_numpy(rng.integers(0, 100, size=P).astype(np.int32)).cuda()
    pb = ops.query("tn_hashgrid_backward_plan_bytes", n, g)
    wb = ops.query("tn_hashgrid_backward_work_bytes", n, g)
    plan = torch.empty(pb, dtype=torch.uint8, device="cuda")
    _capi.call("tn_hashgrid_backward_plan", n, _capi.ptr(info), n, g, _capi.ptr(pos_d), _capi.ptr(plan), pb,
               _capi.stream_ptr())
    ls, lr, b1, b2, eps = 1024.0, 1e-2, 0.9, 0.99, 1e-15
    out = []
    for fused in (False, True):
        w, m, v, st = w0.clone(), m0.clone(), v0.clone(), s0.clone()
        wlp = w0.half() if lp else None
        dt = torch.zeros(P, dtype=torch.float32, device="cuda")
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        args = (0, n, _capi.ptr(info), n, n, g, _capi.ptr(pos_d), _capi.ptr(df_d), _capi.ptr(dt), _capi.ptr(plan), pb,
                _capi.ptr(work), wb)
        if fused:
            a = _capi.Adam(ls, _capi.ptr(info), lr, None, 0.0, b1, b2, eps, _capi.ptr(w), lp, _capi.ptr(wlp),
                           _capi.ptr(m), _capi.ptr(v), _capi.ptr(st))
            _capi.call("tn_hashgrid_backward_planned_adam", *args, ctypes.byref(a), _capi.stream_ptr())
        else:
            _capi.call("tn_hashgrid_backward_planned", *args, _capi.stream_ptr())
            _capi.call("tn_adam_step", P, P, P, ls, _capi.ptr(info), lr, None, 0.0, b1, b2, eps, 0.0, _capi.ptr(w),
                       lp, _capi.ptr(wlp), _capi.ptr(dt), _capi.ptr(m), _capi.ptr(v), _capi.ptr(st),
                       _capi.stream_ptr())
        torch.cuda.synchronize()
        out.append([t.cpu().numpy() for t in (w, m, v, st, dt)] + ([wlp.cpu().numpy()] if lp else []))
    changed = int((out[0][3] != s0.cpu().numpy()).sum())
    print(f"[hash bwd adam lp={lp} nonfinite={nonfinite}] {changed} parameters stepped, "
          f"{int(np.isnan(out[0][0]).sum())} NaN weights", flush=True)
    assert changed > 0.1 * P
    # with the inf, level 0's bin is summed in float by atomics in both runs
    # (summation order, so rounding, differs from run to run): close there,
    # NaNs in the same places (a NaN's sign and payload are not specified),
    # and bit for bit everywhere else
    k0 = 2 * g.level_size[0] if nonfinite else 0
    for a, b, name in zip(out[0], out[1], ("w", "m", "v", "steps", "grad", "w_lp")):
        a0, b0, a, b = a[:k0], b[:k0], a[k0:], b[k0:]
        if k0 and a0.dtype.kind == "f":
            assert np.array_equal(np.isnan(a0), np.isnan(b0)), name
            ok = ~np.isnan(a0)
            assert np.allclose(a0[ok].astype(np.float64), b0[ok], rtol=1e-4, atol=1e-9), name
        elif k0:
            assert np.array_equal(a0, b0), name
        view = np.uint16 if a.dtype == np.float16 else np.uint32
        assert np.array_equal(a.view(view), b.view(view)), name
    assert not out[1][4].any() and not out[0][4].any()


def test_binned_vs_fp64_central_differences(cuda):
    """The trainer's backward (plan, then the gradient passes) against FP64
    central differences, h = 1e-6, of L(table) = sum(dfeat * features) with
    the oracle's fp64 encoding, at 64 touched entries of every level, on
    samples along 64 rays with geometric(48) samples each (the compositing
    and distortion tests' batch shape). Norm-wise 1e-5 per level."""
    import torch

    import oracle
    from turbonerf_amd import ops

    rng = np.random.default_rng(64)
    g = ops.hashgrid_layout()
    steps = np.clip(rng.geometric(1 / 48, 64), 1, 1024)
    pts = []
    for k in steps:
        o = rng.normal(size=3)
        o = 0.5 + 1.2 * o / np.linalg.norm(o)
        d = rng.uniform(0.3, 0.7, size=3) - o
        ln = np.linalg.norm(d)
        t = (ln - 0.45) + (np.sqrt(3.0) / 1024) * np.arange(k)
        pts.append(np.clip(o[None, :] + t[:, None] * (d / ln)[None, :], 0.0, 1.0))
    pos = np.concatenate(pts).astype(np.float32).T.copy()  # [3][n]
    n = pos.shape[1]
    table = rng.uniform(-1e-1, 1e-1, size=g.n_params)
    df = rng.normal(size=(32, n)).astype(np.float32)
    pos_d = torch.from_numpy(pos).cuda().view(-1)
    got = torch.zeros(g.n_params, dtype=torch.float32, device="cuda")
    ops.hashgrid_backward(g, n, pos_d, n, torch.from_numpy(df).cuda().view(-1), got, planned=True)
    got = got.cpu().numpy().astype(np.float64)
    lib = oracle.lib()
    df64 = df.astype(np.float64).ravel()
    feat = np.zeros(32 * n)

    def loss(t):
        lib.oc_hashgrid_fwd_f64(ctypes.addressof(g), n, n, pos.ravel(), t, feat)
        return float(feat @ df64)

    h = 1e-6
    worst = 0.0
    for l in range(g.n_levels):
        lo, hi = 2 * g.level_offset[l], 2 * (g.level_offset[l] + g.level_size[l])
        touched = lo + np.flatnonzero(got[lo:hi])
        pick = rng.choice(touched, min(64, len(touched)), replace=False)
        fd = []
        for k in pick:
            t0 = table[k]
            table[k] = t0 + h
            lp = loss(table)
            table[k] = t0 - h
            lm = loss(table)
            table[k] = t0
            fd.append((lp - lm) / (2 * h))
        e = norm_rel(got[pick], np.array(fd))
        worst = max(worst, e)
        assert e < 1e-5, (l, e)
    print(f"[hash bwd fd] {n} samples on 64 rays, worst level rel err {worst:.2e}", flush=True)
