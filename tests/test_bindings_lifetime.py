"""Object lifetimes across the Python binding (host only, no GPU): a NeRF
returned by NeRFManager.create keeps its manager alive, and a Trainer or a
Renderable keeps its NeRF alive, so dropping the Python names of the manager
and the NeRF never leaves a trainer pointing at freed storage."""
import gc
import weakref


def test_trainer_and_renderable_keep_their_nerf_and_manager_alive():
    from turbonerf_amd import PyTurboNeRF as tn

    for make in (lambda nerf: tn.Trainer(nerf), lambda nerf: tn.Renderable(nerf)):
        mgr = tn.NeRFManager()
        w_mgr = weakref.ref(mgr)
        nerf = mgr.create()
        w_nerf = weakref.ref(nerf)
        holder = make(nerf)
        del mgr, nerf
        gc.collect()
        assert w_mgr() is not None and w_nerf() is not None
        del holder
        gc.collect()
        assert w_mgr() is None and w_nerf() is None
