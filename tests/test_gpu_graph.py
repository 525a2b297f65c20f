"""Training steps replayed from captured HIP graphs (Trainer.train_steps(...,
graph=True): one graph launch per step, the next batch generated inside it,
the learning-rate step and the uniform stream's position read on the
device) against the same steps enqueued eagerly: two trainers with the same
seed and data, one eager and one replaying graphs, hold the same parameters,
Adam state and loss bit for bit after every round of steps, across an
occupancy update (the batch generated ahead is discarded and regenerated
outside the graph) and a change of settings (distortion loss on, alpha
selection on: the graphs are captured again), with eager steps of the graph
trainer between replays."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene(cuda):
    from turbonerf_amd import synthetic

    return synthetic.blob_dataset(seed=21, n_cams=16, width=96, height=96)


def _trainer(scene, precision):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import synthetic

    blobs, cams, images = scene
    mgr = tn.NeRFManager()
    nerf = mgr.create()
    nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
    cfg = tn.TrainingConfig()
    cfg.network.precision = precision
    cfg.network.beta2 = 0.99
    tr = tn.Trainer(nerf, cfg)
    tr.setup_data(batch_size=1 << 17)
    tr.load_images()
    return mgr, nerf, tr


def _same(a, b, tag):
    pa, pb = a.get_params(), b.get_params()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), tag
    sa, sb = a._get_optimizer_state(), b._get_optimizer_state()
    for k in ("m1", "m2", "steps"):
        assert np.array_equal(sa[k].view(np.uint32), sb[k].view(np.uint32)), (tag, k)


@pytest.mark.parametrize("precision", [1, 0])
def test_graph_steps_equal_eager_steps(scene, precision):
    ma, na, eager = _trainer(scene, precision)
    mb, nb, graph = _trainer(scene, precision)
    steps = 0
    for r in range(6):
        if r == 3:  # settings change: both trainers, the graphs are captured again
            for t in (eager, graph):
                t.use_distortion_loss = True
                t.alpha_selection_threshold = 0.5
                t.alpha_selection_probability = 0.5
        me = eager.train_steps(8, read_metrics=True, graph=False)
        mg = graph.train_steps(8, read_metrics=True, graph=True)
        print(f"[graph prec={precision}] round {r}: step {mg.step} loss eager {me.loss:.6f} graph {mg.loss:.6f} "
              f"samples {mg.n_samples}, captures {graph._graph_captures()}", flush=True)
        steps += 8
        assert me.step == mg.step == steps and nb.training_step == na.training_step == steps
        assert np.float32(me.loss) == np.float32(mg.loss) and me.n_samples == mg.n_samples
        _same(eager, graph, f"round {r}")
        if r == 1:  # an occupancy update between replays
            assert eager.update_occupancy_grid(16).n_occupied == graph.update_occupancy_grid(16).n_occupied
        if r == 4:  # an eager step of the graph trainer between replays
            eager.train_step()
            graph.train_step()
            steps += 1
            _same(eager, graph, "eager step between replays")
    assert graph._graph_captures() == 4 and eager._graph_captures() == 0
    assert graph.samples_trained() == eager.samples_trained() > 0
    eager.teardown()
    graph.teardown()
