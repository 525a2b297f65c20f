"""Several NeRFs trained side by side (MultiTrainer: one training step of every
NeRF per round as replays of captured HIP graphs -- each NeRF's own step
graphs on its own streams, or with one_launch one launch of a graph holding
every NeRF's step -- so the NeRFs' steps may run concurrently) against the
same NeRFs stepped one after another with eager steps, as the reference's
controllers are (one NeRFTrainingController per NeRF, src/controllers/
nerf-training-controller.cu:212-260). Three scenes of different seeds; after
every round the parameters, Adam state and losses agree bit for bit, across
an occupancy update of one NeRF, eager steps of one NeRF between rounds (the
NeRFs then sit at different steps and batch slots: a one-launch graph is
captured again) and a settings change of one NeRF (its graphs are captured
again)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = (31, 32, 33)


def _trainer(seed, precision):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = synthetic.blob_dataset(seed=seed, n_cams=12, width=80, height=80)
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
    cfg = tn.TrainingConfig()
    cfg.network.precision = precision
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 16)
    tr.load_images()
    return mgr, nerf, tr


def _same(a, b, tag):
    pa, pb = a.get_params(), b.get_params()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), tag
    sa, sb = a._get_optimizer_state(), b._get_optimizer_state()
    for k in ("m1", "m2", "steps"):
        assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), (tag, k)


@pytest.mark.parametrize("precision,one_launch", [(1, False), (0, False), (1, True)])
def test_multi_nerf_graph_rounds_equal_eager_steps(cuda, precision, one_launch):
    from turbonerf_amd import PyTurboNeRF as tn

    eager = [_trainer(s, precision) for s in SEEDS]
    multi = [_trainer(s, precision) for s in SEEDS]
    mt = tn.MultiTrainer([t for _, _, t in multi], one_launch=one_launch)
    assert mt.one_launch == one_launch
    for r in range(6):
        if r == 4:  # a settings change of one NeRF
            for _, _, t in (eager[2], multi[2]):
                t.use_distortion_loss = True
        for _, _, t in eager:
            t.train_steps(4, read_metrics=False, graph=False)
        mt.train_steps(4)
        mt.synchronize()
        for k, ((_, ne, te), (_, nm, tm)) in enumerate(zip(eager, multi)):
            me, mm = te.train_steps(0, read_metrics=True), tm.train_steps(0, read_metrics=True)
            print(f"[multi prec={precision}] round {r} NeRF {k}: step {nm.training_step} loss eager {me.loss:.6f} "
                  f"graph {mm.loss:.6f}, captures {mt._graph_captures}", flush=True)
            assert ne.training_step == nm.training_step
            assert np.float32(me.loss) == np.float32(mm.loss) and me.n_samples == mm.n_samples
            _same(te, tm, f"round {r} NeRF {k}")
        if r == 1:  # an occupancy update of one NeRF between rounds
            step = multi[1][1].training_step
            assert eager[1][2].update_occupancy_grid(step).n_occupied == multi[1][2].update_occupancy_grid(
                step).n_occupied
        if r == 2:  # an eager step of one NeRF: the NeRFs now use different batch slots
            eager[0][2].train_step()
            multi[0][2].train_step()
            _same(eager[0][2], multi[0][2], "eager step between rounds")
    if one_launch:
        # round graphs for both slots in the first round, after the slot shift
        # and after the settings change
        assert mt._graph_captures == 6
    else:
        # each NeRF's step graphs for both slots, and the changed NeRF's again
        assert mt._graph_captures == 2 * len(SEEDS) + 2
    for (_, _, te), (_, _, tm) in zip(eager, multi):
        assert te.samples_trained() == tm.samples_trained() > 0
    del mt
    for _, _, t in eager + multi:
        t.teardown()
