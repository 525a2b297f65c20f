"""ctypes binding of the ``tn_*`` C ABI (include/turbonerf/tn_c_api.h).

The kernel library must have been built (``python -m turbonerf_amd.build``);
importing this module on a machine without it raises immediately -- there is
no CPU fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libturbonerf_kernels.so"

TN_STATUS = {
    0: "TN_OK",
    1: "TN_ERR_INVALID_ARGUMENT",
    2: "TN_ERR_LAUNCH",
    3: "TN_ERR_UNSUPPORTED",
    4: "TN_ERR_OUT_OF_MEMORY",
    5: "TN_ERR_NO_RAYS",
    6: "TN_ERR_NO_SAMPLES",
}


class TNError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn} failed: {TN_STATUS.get(code, code)}")
        self.code = code


# ----------------------------------------------------------------- structs
class BBox(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("min_x", "min_y", "min_z", "max_x", "max_y", "max_z")]

    @classmethod
    def cube(cls, size: float) -> "BBox":
        h = 0.5 * size
        return cls(-h, -h, -h, h, h, h)


class Camera(C.Structure):
    _fields_ = [
        ("resolution", C.c_int32 * 2),
        ("near_plane", C.c_float),
        ("far_plane", C.c_float),
        ("focal_length", C.c_float * 2),
        ("resolution_f", C.c_float * 2),
        ("principal_point", C.c_float * 2),
        ("shift", C.c_float * 2),
        ("offset_px", C.c_float * 2),
        ("transform", C.c_float * 12),
        ("dist_params", C.c_float * 5),
        ("show_image_planes", C.c_int32),
    ]


class OccupancyGrid(C.Structure):
    _fields_ = [
        ("bitfield", C.c_void_p),
        ("density", C.c_void_p),
        ("n_levels", C.c_int32),
        ("resolution", C.c_int32),
        ("level_dt_scale", C.c_float),
    ]


MAX_LEVELS = 16


class HashGrid(C.Structure):
    _fields_ = [
        ("n_levels", C.c_int32),
        ("n_features", C.c_int32),
        ("log2_hashmap_size", C.c_int32),
        ("base_resolution", C.c_int32),
        ("per_level_scale", C.c_float),
        ("n_params", C.c_uint32),
        ("scale", C.c_float * MAX_LEVELS),
        ("resolution", C.c_uint32 * MAX_LEVELS),
        ("level_size", C.c_uint32 * MAX_LEVELS),
        ("level_offset", C.c_uint32 * MAX_LEVELS),
        ("level_dense", C.c_uint32 * MAX_LEVELS),
    ]


# ----------------------------------------------------------------- loading
_lib: C.CDLL | None = None

class RenderNerf(C.Structure):
    _fields_ = [
        ("to_local", C.c_float * 12),
        ("render_box", BBox),
        ("repeat_source", BBox),
        ("repeat", C.c_int32),
        ("inv_aabb", C.c_float),
        ("training_min", C.c_float * 3),
        ("training_size", C.c_float),
    ]


class Appearance(C.Structure):
    _fields_ = [
        ("bias", C.c_void_p),
        ("image_id", C.c_void_p),
        ("n_appearances", C.c_uint32),
        ("dz", C.c_void_p),
    ]


class Adam(C.Structure):
    """tn_adam_t: NGP Adam settings for the hash-table range."""
    _fields_ = [
        ("loss_scale", C.c_float),
        ("scale_by_rays_info", C.c_void_p),
        ("learning_rate", C.c_float),
        ("d_step", C.c_void_p),
        ("lr_decay", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("weights", C.c_void_p),
        ("lp_dtype", C.c_int),
        ("weights_lp", C.c_void_p),
        ("first_moments", C.c_void_p),
        ("second_moments", C.c_void_p),
        ("param_steps", C.c_void_p),
    ]


vp = C.c_void_p
u32 = C.c_uint32
u64 = C.c_uint64
i32 = C.c_int32
f32 = C.c_float
sz = C.c_size_t

_SIGS: dict[str, tuple] = {
    "tn_fill_uniform": (C.c_int, [u64, u64, u64, vp, vp]),
    "tn_fill_uniform_dev": (C.c_int, [u64, u64, vp, vp, vp]),
    "tn_init_training_rays": (C.c_int, [u32, vp, u32, u32, u32, i32, i32, f32, vp, vp, vp, vp, vp, vp, vp,
                                        vp, vp, vp, vp, vp]),
    "tn_deactivate_rays": (C.c_int, [u32, vp, u32, f32, f32, vp, vp, vp, vp]),
    "tn_march_count": (C.c_int, [u32, vp, u32, BBox, OccupancyGrid, f32, f32, f32, u32, vp, vp, vp, vp, vp,
                                 vp, vp]),
    "tn_scan_workspace_bytes": (sz, [u32]),
    "tn_scan_select": (C.c_int, [C.c_int, u32, vp, vp, u32, vp, vp, vp, sz, vp]),
    "tn_march_generate": (C.c_int, [u32, vp, u32, u32, BBox, f32, OccupancyGrid, f32, f32, f32, vp, vp, vp, vp, vp,
                                    vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tn_estimate_rays": (C.c_int, [vp, u32, u32, u32, vp, vp]),
    "tn_accumulate_samples": (C.c_int, [vp, vp, vp]),
    "tn_batch_tail": (C.c_int, [vp, u32, vp, vp]),
    "tn_composite_scratch_bytes": (sz, [u32, u32]),
    "tn_composite_train": (C.c_int, [C.c_int, vp, u32, u32, u32, f32, C.c_int, f32, vp, vp, vp, vp, vp, vp, vp, vp,
                                     vp, vp, vp, sz, vp]),
    "tn_composite_train_noalloc": (C.c_int, [C.c_int, vp, u32, u32, u32, f32, C.c_int, f32, vp, vp, vp, vp, vp, vp,
                                             vp, vp, vp, vp, vp]),
    "tn_composite_train_noalloc_max_bytes": (sz, []),
    "tn_composite_forward": (C.c_int, [C.c_int, vp, u32, u32, u32, f32, C.c_int, vp, vp, vp, vp, vp, vp, vp]),
    "tn_reduce_loss": (C.c_int, [vp, u32, u32, vp, vp, vp]),
    "tn_step_metrics": (C.c_int, [vp, u32, vp, vp, vp, vp, vp, vp]),
    "tn_occ_decay": (C.c_int, [OccupancyGrid, f32, vp]),
    "tn_occ_sample_points": (C.c_int, [u32, u32, u32, OccupancyGrid, C.c_int, f32, vp, vp, vp]),
    "tn_occ_update_density": (C.c_int, [C.c_int, u32, u32, OccupancyGrid, C.c_int, C.c_int, vp, vp, vp]),
    "tn_occ_update_bits": (C.c_int, [OccupancyGrid, f32, vp]),
    "tn_popcount_workspace_bytes": (sz, [u64]),
    "tn_popcount": (C.c_int, [vp, u64, vp, vp, sz, vp]),
    "tn_blob_density_grid": (C.c_int, [OccupancyGrid, vp, u32, vp]),
    "tn_blob_eval": (C.c_int, [vp, u32, u32, f32, vp, u32, vp, vp, vp, vp]),
    "tn_hashgrid_layout": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, vp]),
    "tn_hashgrid_forward": (C.c_int, [C.c_int, C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, vp]),
    "tn_hashgrid_backward": (C.c_int, [C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, vp]),
    "tn_hashgrid_backward_workspace_bytes": (C.c_size_t, [u32, HashGrid]),
    "tn_hashgrid_backward_binned": (C.c_int, [C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, vp, C.c_size_t, vp]),
    "tn_hashgrid_backward_plan_bytes": (C.c_size_t, [u32, HashGrid]),
    "tn_hashgrid_backward_work_bytes": (C.c_size_t, [u32, HashGrid]),
    "tn_hashgrid_backward_plan": (C.c_int, [u32, vp, u32, HashGrid, vp, vp, C.c_size_t, vp]),
    "tn_hashgrid_backward_planned": (C.c_int, [C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, vp, C.c_size_t, vp,
                                               C.c_size_t, vp]),
    "tn_hashgrid_backward_part_samples": (u32, [u32, HashGrid]),
    "tn_hashgrid_backward_scatter": (C.c_int, [C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, C.c_size_t, vp,
                                               C.c_size_t, u32, u32, vp]),
    "tn_hashgrid_backward_accumulate": (C.c_int, [u32, HashGrid, vp, vp, C.c_size_t, vp, C.c_size_t, vp]),
    "tn_hashgrid_backward_planned_adam": (C.c_int, [C.c_int, u32, vp, u32, u32, HashGrid, vp, vp, vp, vp,
                                                    C.c_size_t, vp, C.c_size_t, vp, vp]),
    "tn_hashgrid_backward_accumulate_adam": (C.c_int, [u32, HashGrid, vp, vp, C.c_size_t, vp, C.c_size_t, vp, vp]),
    "tn_mlp_forward": (C.c_int, [C.c_int, u32, vp, u32, u32, u32, vp, u32, vp, vp, C.c_int, vp, vp, vp, vp]),
    "tn_mlp_inference": (C.c_int, [C.c_int, C.c_int, u32, vp, u32, u32, u32, vp, u32, vp, vp, C.c_int, vp, vp, vp,
                                   vp]),
    "tn_mlp_backward_workspace_bytes": (sz, [u32]),
    "tn_mlp_backward": (C.c_int, [C.c_int, u32, vp, u32, u32, vp, u32, vp, vp, vp, vp, u32, vp, vp, f32, C.c_int, vp,
                                  vp, C.c_int, vp, sz, vp, vp]),
    "tn_mlp_lds_bytes": (sz, [C.c_int, C.c_int]),
    "tn_adam_step": (C.c_int, [u32, u32, u32, f32, vp, f32, vp, f32, f32, f32, f32, f32, vp, C.c_int, vp, vp, vp, vp,
                               vp, vp]),
    "tn_cast_params": (C.c_int, [u64, vp, C.c_int, vp, vp]),
    "tn_render_rays": (C.c_int, [u32, u32, u32, Camera, BBox, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tn_render_write_rgba": (C.c_int, [u32, vp, u32, vp, vp, vp, vp, vp]),
    "tn_distortion_scratch_bytes": (sz, [u32, u32]),
    "tn_distortion": (C.c_int, [C.c_int, vp, u32, u32, f32, vp, vp, vp, vp, vp, vp, vp, vp, sz, vp]),
    "tn_appearance_bias": (C.c_int, [u32, vp, vp, vp, vp]),
    "tn_appearance_workspace_bytes": (sz, [u32]),
    "tn_appearance_backward": (C.c_int, [u32, vp, u32, vp, vp, u32, vp, vp, vp, vp, vp, sz, vp]),
    "tn_render_rays_multi": (C.c_int, [u32, u32, u32, Camera, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tn_render_march_fixed": (C.c_int, [u32, u32, u32, f32, f32, f32, u32, vp, vp, vp, u32, vp, vp, vp, vp, vp, vp]),
    "tn_render_localize": (C.c_int, [u32, u32, u32, RenderNerf, OccupancyGrid, vp, u32, vp, vp, vp, vp, vp, vp, vp]),
    "tn_render_compact": (C.c_int, [u32, u32, vp, vp, vp, vp, vp, vp, vp]),
    "tn_render_accumulate": (C.c_int, [C.c_int, u32, u32, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]),
    "tn_render_composite_fixed": (C.c_int, [u32, u32, u32, f32, vp, vp, vp, vp, vp, vp, vp, vp]),
    "tn_render_expand_nearest": (C.c_int, [u32, u32, u32, u32, vp, vp, vp]),
    "tn_count_nonzero": (C.c_int, [u32, vp, vp, vp]),
    "tn_fill_u32": (C.c_int, [u32, vp, u32, vp]),
    "tn_tune_set": (C.c_int, [C.c_int, C.c_int]),
    "tn_tune_get": (C.c_int, [C.c_int, vp]),
    "tn_tune_generation": (u32, []),
}

# entry points added by later kernel files register their signatures here
EXTRA_SIGS: dict[str, tuple] = {}


def declared_symbols() -> list[str]:
    return sorted(set(_SIGS) | set(EXTRA_SIGS))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"TurboNeRF kernel library not built: {LIB_PATH} is missing "
                "(run `python -m turbonerf_amd.build`)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in {**_SIGS, **EXTRA_SIGS}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise TNError on failure."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise TNError(name, rc)


def query(name: str, *args) -> int:
    return int(getattr(lib(), name)(*args))


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor, None for None."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
