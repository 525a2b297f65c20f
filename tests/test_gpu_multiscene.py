"""Several scenes at once (the multi-scene configuration): eight NeRFs in one
NeRFManager, hash tables of 2^15 ... 2^21, 2^21 entries per level, seeded blob
fields of 4 to 32 blobs, uneven sample budgets summing to 2^19, trained in
interleaved steps on one stream each. Every scene's loss is finite and falls,
and the largest tables (2^21 entries: 5,717 bins) run the binned hash-grid
backward."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LOG2_T = [15, 16, 17, 18, 19, 20, 21, 21]
BLOBS = [4, 8, 12, 16, 20, 24, 28, 32]
BATCH = [16384, 32768, 65536, 131072, 131072, 65536, 49152, 32768]


def test_eight_scenes_interleaved(cuda):
    from turbonerf_amd import PyTurboNeRF as tn
    from turbonerf_amd import _capi, ops, synthetic

    assert sum(BATCH) == 1 << 19
    mgr = tn.NeRFManager()
    trainers = []
    for k in range(8):
        blobs, cams, images = synthetic.blob_dataset(seed=100 + k, n_cams=8, width=64, height=64, n_blobs=BLOBS[k])
        nerf = mgr.create()
        nerf.attach_dataset(tn.Dataset(tn.BoundingBox(1.0), synthetic.to_pyturbonerf_cameras(tn, cams), images))
        cfg = tn.TrainingConfig()
        cfg.network.precision = 1
        cfg.network.log2_hashmap_size = LOG2_T[k]
        cfg.network.beta2 = 0.99
        cfg.rng_seed = 7 + k
        tr = tn.Trainer(nerf, cfg)
        tr.setup_data(batch_size=BATCH[k])
        tr.load_images()
        trainers.append(tr)
        g = ops.hashgrid_layout(log2_T=LOG2_T[k])
        assert _capi.query("tn_hashgrid_backward_workspace_bytes", BATCH[k], g) > 0
    losses = [[] for _ in trainers]
    for i in range(80):
        for k, tr in enumerate(trainers):
            m = tr.train_step()
            assert np.isfinite(m.loss) and 0 < m.n_samples <= BATCH[k]
            losses[k].append(m.loss)
    for k in range(8):
        first, last = np.mean(losses[k][:8]), np.mean(losses[k][-8:])
        print(f"[multiscene] scene {k} T=2^{LOG2_T[k]} blobs {BLOBS[k]} batch {BATCH[k]}: loss {first:.4f} -> "
              f"{last:.4f}", flush=True)
        assert last < 0.7 * first
