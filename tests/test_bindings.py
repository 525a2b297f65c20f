"""The Python binding on the host (no GPU): the classes and members the
reference module exposes for training and rendering
(src/api/python_bindings.cu:56-693; the Blender bridge and OpenGL/CUDA render
surfaces are not part of this build, and the runtime checks name HIP and
gfx950 instead of CUDA versions and compute capabilities), the
UpdatableProperty wrappers, and object lifetimes: a NeRF returned by
NeRFManager.create keeps its manager alive, and a Trainer or a Renderable
keeps its NeRF alive, so dropping the Python names of the manager and the
NeRF never leaves a trainer pointing at freed storage."""
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


SURFACE = {
    "Camera": ["resolution", "near", "far", "focal_length", "principal_point", "shift", "transform",
               "dist_params", "show_image_planes"],
    "DistortionParams": ["k1", "k2", "k3", "p1", "p2"],
    "Dataset": ["load_transforms", "copy", "to_json", "set_camera_at", "file_path", "cameras", "bounding_box",
                "images", "image_dimensions", "contains_multiple_image_dims"],
    "TrainingImage": ["file_path", "dimensions"],
    "BoundingBox": ["size", "max_x", "max_y", "max_z", "min_x", "min_y", "min_z"],
    "Transform4f": ["from_nerf", "to_nerf", "inverse", "to_matrix"],
    "NeRF": ["attach_dataset", "detach_dataset", "is_dirty", "can_train", "is_image_data_loaded", "render_bbox",
             "training_bbox", "transform", "is_visible", "is_dataset_dirty", "dataset", "can_render",
             "training_step", "id"],
    "NeRFManager": ["create", "destroy"],
    "Trainer": ["setup_data", "teardown", "reset_training", "load_images", "train_step", "update_occupancy_grid",
                "alpha_selection_threshold", "alpha_selection_probability", "min_step_size", "use_distortion_loss"],
    "Renderer": ["submit", "cancel"],
    "RenderRequest": ["cancel"],
    "Renderable": ["nerf", "bounding_box"],
    "CPURenderBuffer": ["get_rgba"],
    "RenderTarget": ["save_image", "set_size", "width", "height", "free"],
    "RenderModifiers": [], "RenderFlags": [], "RenderPattern": [],
    "FileManager": ["load", "save"],
    "RuntimeManager": ["gpu_archs", "required_runtime_version", "compiled_runtime_version", "current_runtime_version",
                       "is_driver_version_sufficient", "get_device_architecture", "check_runtime"],
    "RuntimeVersion": ["major", "minor", "patch", "to_string"],
    "UpdatableBoundingBox": ["get", "set", "set_dirty", "is_dirty"],
    "UpdatableTransform4f": ["get", "set", "set_dirty", "is_dirty"],
}


def test_reference_surface_is_exposed():
    from turbonerf_amd import PyTurboNeRF as tn

    assert tn.__version__ == "0.0.19" and callable(tn.teardown)
    missing = [c for c in SURFACE if not hasattr(tn, c)]
    missing += [f"{c}.{a}" for c, attrs in SURFACE.items() for a in attrs
                if hasattr(tn, c) and not hasattr(getattr(tn, c), a)]
    assert not missing, missing
    u = tn.UpdatableTransform4f()
    assert u.is_dirty()
    u.set(tn.Transform4f.Identity())
    u.set_dirty(False)
    assert not u.is_dirty() and u.get() == tn.Transform4f.Identity()


def test_destroyed_slot_is_reused_in_place():
    """NeRFManager.destroy then create reuses the slot: the old handle stays
    valid and refers to the new NeRF (no dangling pointer)."""
    from turbonerf_amd import PyTurboNeRF as tn

    mgr = tn.NeRFManager()
    a = mgr.create()
    slot = a.id
    mgr.destroy(a)
    b = mgr.create()
    assert b.id == slot and a.id == slot and a.training_step == 0 and not a.can_render


def test_runtime_manager_names_the_build():
    """The build's architecture list, the HIP version it was compiled with
    (the installed hip_version.h) and the runtime it requires (that major
    version)."""
    import re
    from pathlib import Path

    from turbonerf_amd import PyTurboNeRF as tn

    rm = tn.RuntimeManager()
    assert list(rm.gpu_archs) == ["gfx950"]
    hdr = Path("/opt/rocm/include/hip/hip_version.h")
    if hdr.exists():
        v = {k: int(re.search(rf"#define HIP_VERSION_{k} (\d+)", hdr.read_text()).group(1))
             for k in ("MAJOR", "MINOR", "PATCH")}
        c = tn.RuntimeManager.compiled_runtime_version()
        assert (c.major, c.minor, c.patch) == (v["MAJOR"], v["MINOR"], v["PATCH"])
        assert str(c) == f"{v['MAJOR']}.{v['MINOR']}.{v['PATCH']}"
        r = tn.RuntimeManager.required_runtime_version()
        assert (r.major, r.minor, r.patch) == (v["MAJOR"], 0, 0)
