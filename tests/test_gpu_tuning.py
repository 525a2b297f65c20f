"""Launch-shape knobs and the autotuner on the GPU.

* Every knob but the MLP backward's workgroups per CU and the two-half step
  leaves training results bit for bit unchanged (among them Adam on the hash
  table fused into the hash-grid backward, against its own pass) (the hash-grid backward's workgroups per
  level shape only the one-call binned backward, not the trainer's planned
  one): two trainers with the same seed and data,
  one stepped under the default launch shapes and one under other values of
  every knob, hold the same parameters and Adam state after each round of
  eager steps, and after captured steps whose graphs are captured again
  when a knob changes.
* Two MLP backward workgroups per CU only reorder the fp32 sum of the
  per-workgroup weight-gradient partials: the parameters after one step
  agree to within a few fp32 ulps of the update. The same holds for a step
  over the batch's two halves on two streams, whose hash-grid update is
  the whole batch's bit for bit.
* Trainer.autotune and MultiTrainer.autotune return a pick from each
  searched knob's candidates with a timing for every candidate, and
  training goes on.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEFAULTS = {"hash_fwd_block": 256, "hash_bwd_level_blocks": 128, "mlp_fwd_blocks_per_cu": 3,
            "mlp_bwd_blocks_per_cu": 1, "march_max_blocks": 2048, "adam_block": 256, "hash_bwd_level_groups": 1,
            "hash_fwd_lds_levels": 0, "step_halves": 0, "hash_adam_fused": 0}
OTHER = {"hash_fwd_block": 64, "hash_bwd_level_blocks": 16, "mlp_fwd_blocks_per_cu": 1, "march_max_blocks": 512,
         "adam_block": 1024, "hash_bwd_level_groups": 3, "hash_fwd_lds_levels": 2, "step_halves": 0,
         "hash_adam_fused": 1}


def _set(tn, knobs):
    for k, v in knobs.items():
        tn.set_tuning(k, v)


@pytest.fixture
def tuning(cuda):
    from turbonerf_amd import PyTurboNeRF as tn

    assert tn.get_tuning() == DEFAULTS
    yield tn
    _set(tn, DEFAULTS)


def _trainer(seed, precision=1, batch=1 << 16):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=seed, n_cams=12, width=80, height=80)
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
    cfg = tn.TrainingConfig()
    cfg.network.precision = precision
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=batch)
    tr.load_images()
    return mgr, nerf, tr


def _same(a, b, tag):
    assert np.array_equal(a.get_params().view(np.uint32), b.get_params().view(np.uint32)), tag
    sa, sb = a._get_optimizer_state(), b._get_optimizer_state()
    for k in ("m1", "m2", "steps"):
        assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), (tag, k)


@pytest.mark.parametrize("precision", [1, 0])
def test_launch_shapes_leave_results_unchanged(tuning, precision):
    tn = tuning
    ref, alt = _trainer(41, precision), _trainer(41, precision)
    for r in range(3):
        _set(tn, DEFAULTS)
        mr = ref[2].train_steps(4, read_metrics=True, graph=False)
        _set(tn, OTHER)
        ma = alt[2].train_steps(4, read_metrics=True, graph=False)
        print(f"[tuning prec={precision}] eager round {r}: loss {mr.loss:.6f} / {ma.loss:.6f}", flush=True)
        assert np.float32(mr.loss) == np.float32(ma.loss)
        _same(ref[2], alt[2], f"eager round {r}")
    caps = alt[2]._graph_captures()
    for r, knobs in enumerate(({"march_max_blocks": 1024}, {"adam_block": 128}, {"hash_bwd_level_groups": 16},
                               {"hash_adam_fused": 1, "hash_bwd_level_groups": 4})):
        _set(tn, DEFAULTS)
        ref[2].train_steps(4, read_metrics=False, graph=True)
        _set(tn, {**OTHER, **knobs})
        alt[2].train_steps(4, read_metrics=False, graph=True)
        _same(ref[2], alt[2], f"graph round {r}")
    # the alternating knobs made each of the graph trainer's replays a new capture set
    assert alt[2]._graph_captures() >= caps + 6
    for t in (ref, alt):
        t[2].teardown()


@pytest.mark.parametrize("precision", [1, 0])
def test_step_halves_reorder_the_weight_gradient_sums_only(tuning, precision):
    """A step over the batch's two halves on two streams: every per-sample
    result is the whole batch's, so the hash-grid gradient, and with it the
    table's Adam update, is the same bit for bit; the MLP weight gradient is
    the first half's sum plus the second's, a reordered fp32 sum. Then eager
    and captured steps of the split trainer go on to agree bit for bit."""
    tn = tuning
    ref, alt = _trainer(47, precision), _trainer(47, precision)
    tn.set_tuning("step_halves", 0)
    mr = ref[2].train_steps(1, read_metrics=True, graph=False)
    tn.set_tuning("step_halves", 1)
    ma = alt[2].train_steps(1, read_metrics=True, graph=False)
    p0, p1 = ref[2].get_params(), alt[2].get_params()
    n_grid = len(p0) - 9408
    d = np.abs(p1 - p0)
    print(f"[tuning prec={precision}] step halves: loss {mr.loss:.6f} / {ma.loss:.6f}, network max |dp| "
          f"{d[n_grid:].max():.2e}", flush=True)
    assert np.float32(mr.loss) == np.float32(ma.loss)
    assert np.array_equal(p0[:n_grid].view(np.uint32), p1[:n_grid].view(np.uint32))
    assert d[n_grid:].max() < 1e-4
    # split steps, eager against captured
    eag, cap = _trainer(48, precision), _trainer(48, precision)
    for r in range(3):
        eag[2].train_steps(4, read_metrics=False, graph=False)
        cap[2].train_steps(4, read_metrics=False, graph=True)
        _same(eag[2], cap[2], f"halves round {r}")
    # split steps with the table's Adam fused into the sum pass: the same bits
    unf, fus = _trainer(49, precision), _trainer(49, precision)
    for r in range(2):
        tn.set_tuning("hash_adam_fused", 0)
        unf[2].train_steps(3, read_metrics=False, graph=r == 1)
        tn.set_tuning("hash_adam_fused", 1)
        fus[2].train_steps(3, read_metrics=False, graph=r == 1)
        _same(unf[2], fus[2], f"halves fused round {r}")
    for t in (ref, alt, eag, cap, unf, fus):
        t[2].teardown()


def test_two_mlp_backward_blocks_per_cu_reorders_the_partial_sums_only(tuning):
    tn = tuning
    ref, alt = _trainer(43), _trainer(43)
    _set(tn, DEFAULTS)
    ref[2].train_step()
    tn.set_tuning("mlp_bwd_blocks_per_cu", 2)
    alt[2].train_step()
    p0, p1 = ref[2].get_params(), alt[2].get_params()
    n_grid = len(p0) - 9408
    d = np.abs(p1 - p0)
    print(f"[tuning] mlp bwd 2 blocks/CU: hash max |dp| {d[:n_grid].max():.2e}, network max |dp| "
          f"{d[n_grid:].max():.2e}", flush=True)
    # Adam's first step moves each weight by lr * g / (|g| + eps) ~ lr: a
    # reordered fp32 sum changes a weight's gradient by a few ulps and so
    # its update by far less than lr; hash entries see the same dfeat up to
    # the same rounding
    assert d.max() < 1e-4
    assert np.array_equal(ref[2]._get_optimizer_state()["steps"], alt[2]._get_optimizer_state()["steps"])
    for t in (ref, alt):
        t[2].teardown()


def test_autotune_picks_from_the_candidates_and_training_goes_on(tuning):
    tn = tuning
    single = _trainer(45)
    single[2].train_steps(16, read_metrics=False, graph=True)
    res = single[2].autotune(rounds_per_trial=4, repetitions=1)
    print(f"[tuning] single: {res['picks']} {res['ns_per_sample_before']:.3f} -> "
          f"{res['ns_per_sample_after']:.3f} ns/sample", flush=True)
    # the hash-grid backward's workgroups per level are not searched (the
    # trainer's backward plans use a fixed 256), the other nine are
    cands = {"hash_fwd_block": [64, 128, 256], "mlp_fwd_blocks_per_cu": [1, 2, 3, 4], "mlp_bwd_blocks_per_cu": [1, 2],
             "march_max_blocks": [512, 1024, 2048, 4096], "adam_block": [128, 256, 512, 1024],
             "hash_bwd_level_groups": [1, 2, 4, 8], "hash_fwd_lds_levels": [0, 1, 2], "step_halves": [0, 1],
             "hash_adam_fused": [0, 1]}
    assert set(res["picks"]) == set(DEFAULTS) and tn.get_tuning() == res["picks"]
    assert set(res["trials"]) == set(cands)
    for k, c in cands.items():
        assert res["picks"][k] in c and [v for v, _ in res["trials"][k]] == c
        assert all(t > 0 for _, t in res["trials"][k])
    m = single[2].train_step()
    assert np.isfinite(m.loss) and m.n_samples > 0
    _set(tn, DEFAULTS)
    multi = [_trainer(46 + k, batch=1 << 15) for k in range(3)]
    mt = tn.MultiTrainer([t for _, _, t in multi])
    mt.train_steps(8)
    s0 = mt.samples_trained()
    res = mt.autotune(rounds_per_trial=2, repetitions=1)
    print(f"[tuning] multi: {res['picks']}", flush=True)
    assert all(res["picks"][k] in c for k, c in cands.items())
    assert mt.samples_trained() > s0
    del mt
    for t in [single] + multi:
        t[2].teardown()
