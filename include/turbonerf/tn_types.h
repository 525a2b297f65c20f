/*
 * Plain-C data formats shared by the TurboNeRF MI355X kernel library, its C++
 * host layer and any FFI binding. Every struct is a POD that is passed by
 * value (or by device pointer) to the `tn_*` entry points declared in
 * tn_c_api.h. Layouts follow the reference's device-copyable PODs field by
 * field where one exists, so a host that already fills the reference structs
 * can fill these without reinterpretation.
 */
#ifndef TURBONERF_TN_TYPES_H
#define TURBONERF_TN_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point (0 = success). */
enum tn_status {
    TN_OK = 0,
    TN_ERR_INVALID_ARGUMENT = 1,   /* null pointer, zero/oversized count, bad stride */
    TN_ERR_LAUNCH = 2,             /* hipLaunchKernel / hipGetLastError failure */
    TN_ERR_UNSUPPORTED = 3,        /* feature outside the supported configuration */
    TN_ERR_OUT_OF_MEMORY = 4,      /* caller-provided workspace too small */
    TN_ERR_NO_RAYS = 5,            /* "No rays were generated" (reference trainer.cu:162-164) */
    TN_ERR_NO_SAMPLES = 6          /* "No samples were generated" (reference trainer.cu:181-183) */
};

/* Launch-shape knobs of the kernel launchers (tn_tune_set / tn_tune_get):
 * process-wide, read when a kernel is enqueued. Every value gives the same
 * results except TN_TUNE_MLP_BWD_BLOCKS_PER_CU and TN_TUNE_STEP_HALVES, which
 * change the order in which the MLP weight gradient's per-workgroup partial
 * sums are added (fp32 rounding only). Defaults in parentheses. */
enum tn_tune_knob {
    TN_TUNE_HASH_FWD_BLOCK = 0,        /* threads per hash-forward workgroup: 64, 128 or 256 (256) */
    TN_TUNE_HASH_BWD_LEVEL_BLOCKS = 1, /* binned hash backward: count/scatter workgroups per level, 8..256 (128) */
    TN_TUNE_MLP_FWD_BLOCKS_PER_CU = 2, /* MLP forward workgroups per compute unit, 1..4 (3) */
    TN_TUNE_MLP_BWD_BLOCKS_PER_CU = 3, /* MLP backward workgroups per compute unit, 1..2 (1) */
    TN_TUNE_MARCH_MAX_BLOCKS = 4,      /* ray-marching workgroups at most, 256..8192 (2048) */
    TN_TUNE_ADAM_BLOCK = 5,            /* threads per Adam workgroup, 64..1024, a multiple of 64 (256) */
    TN_TUNE_HASH_BWD_LEVEL_GROUPS = 6, /* binned hash backward: levels scattered and summed in this many groups, 1..16 (1) */
    TN_TUNE_HASH_FWD_LDS_LEVELS = 7,   /* hash forward: coarse dense levels read from a copy in LDS, at most, 0..4 (0);
                                          only those whose tables fit 64 KiB together are staged */
    TN_TUNE_STEP_HALVES = 8,           /* training step: the per-sample kernels over the batch's two halves on two
                                          streams, so one half's MLP passes overlap the other's hash-grid passes, 0..1 (0) */
    TN_TUNE_HASH_ADAM_FUSED = 9,       /* training step: Adam on the hash table applied by the hash-grid backward's
                                          sum passes instead of a pass of its own, 0..1 (0) */
    TN_TUNE_COUNT = 10
};

/* Axis-aligned box; same field order as BoundingBox
 * (reference src/models/bounding-box.cuh:8-24). */
typedef struct tn_bbox {
    float min_x, min_y, min_z;
    float max_x, max_y, max_z;
} tn_bbox_t;

/* Camera; fields mirror Camera (reference src/models/camera.cuh:45-79).
 * transform is the 3x4 camera-to-world matrix, row-major
 * (m00 m01 m02 m03, m10 ..., m20 ...), as Transform4f stores it
 * (reference src/math/transform4f.cuh). offset_px = shift*resolution - principal_point. */
typedef struct tn_camera {
    int32_t resolution[2];
    float near_plane;
    float far_plane;
    float focal_length[2];
    float resolution_f[2];
    float principal_point[2];
    float shift[2];
    float offset_px[2];
    float transform[12];
    float dist_params[5];      /* k1 k2 k3 p1 p2 lens model; all zero: pinhole */
    int32_t show_image_planes;
} tn_camera_t;

/* Multi-cascade occupancy grid (reference src/core/occupancy-grid.cuh:22-45).
 * bitfield: resolution^3 bytes in Morton order; bit k of a byte = cascade k.
 * density: n_levels * resolution^3 floats, cascade-major.
 * level_dt_scale: the marching kernels choose a sample's cascade from
 * dt * level_dt_scale. Training marches with the world step (1, reference
 * training-batch-kernels.cuh:273,441); the reference's renderer passes the
 * step normalised by the scene size (1/aabb, rendering-kernels.cu:445,447).
 * 0 reads as 1. */
typedef struct tn_occupancy_grid {
    uint8_t* bitfield;
    float* density;
    int32_t n_levels;
    int32_t resolution;
    float level_dt_scale;
} tn_occupancy_grid_t;

/* Multiresolution hash-grid geometry (Instant-NGP encoding configured at
 * reference src/core/nerf-network.cu:86-104). Filled by tn_hashgrid_layout(). */
#define TN_HASHGRID_MAX_LEVELS 16
typedef struct tn_hashgrid {
    int32_t n_levels;              /* L (16) */
    int32_t n_features;            /* F (2) */
    int32_t log2_hashmap_size;     /* log2 T (19) */
    int32_t base_resolution;       /* N_min (16) */
    float per_level_scale;         /* b */
    uint32_t n_params;             /* sum of level sizes * F */
    float scale[TN_HASHGRID_MAX_LEVELS];          /* N_min * b^l - 1 */
    uint32_t resolution[TN_HASHGRID_MAX_LEVELS];  /* ceil(scale) + 1 */
    uint32_t level_size[TN_HASHGRID_MAX_LEVELS];  /* entries in level l */
    uint32_t level_offset[TN_HASHGRID_MAX_LEVELS];/* first entry of level l */
    uint32_t level_dense[TN_HASHGRID_MAX_LEVELS]; /* 1 = dense index, 0 = hashed */
} tn_hashgrid_t;

/* Parameter layout of one NeRF's network (flat buffer, hash grid first):
 *   [0, n_grid)                        hash-grid table, level-major, F features per entry
 *   [n_grid, n_grid + 3072)            density MLP: W1[64][32], W2[16][64]
 *   [n_grid + 3072, n_grid + 3072+6336) colour MLP: W3[64][32], W4[64][64], W5[3][64]
 * All matrices are row-major [out][in]; the MLPs have no biases. */
#define TN_DENSITY_MLP_PARAMS (64 * 32 + 16 * 64)
#define TN_COLOR_MLP_PARAMS (64 * 32 + 64 * 64 + 3 * 64)

/* One NeRF of a render of several (reference NeRFRenderable with its
 * spatial effects, src/models/nerf-renderable.cuh, src/effects/spatial/
 * repeater.cu:8-76): the world -> NeRF transform (3x4 row-major), the box a
 * sample must fall in, in the NeRF frame (the render box grown by the
 * effects), and an optional repeater that folds positions into its source
 * box (p -> ((p - min) mod size + size) mod size + min) before the grid and
 * the network see them. */
typedef struct tn_render_nerf {
    float to_local[12];
    tn_bbox_t render_box;
    tn_bbox_t repeat_source;
    int32_t repeat;          /* 1: fold positions into repeat_source */
    float inv_aabb;          /* 1 / the NeRF's training-box side (step normalisation) */
    float training_min[3];   /* training box: network inputs are (p - min) / size */
    float training_size;
} tn_render_nerf_t;

/* Appearance embedding (reference src/core/embedding.cuh:45-99,
 * src/core/nerf-network.cu:27-72): with it the colour MLP's first layer takes
 * 16 more inputs, the embedding emb[k][16] of the sample's image k, through
 * 16 more weight columns W3e[64][16]. Their contribution W3e emb[k] is the
 * same for every sample of image k, so it enters the MLP kernels as a
 * per-image bias of that layer (tn_appearance_bias), and the gradients of
 * W3e and emb follow from the per-image sums of that layer's pre-activation
 * gradient dz (tn_appearance_backward). */
#define TN_APPEARANCE_DIMS 16
#define TN_APPEARANCE_WEIGHTS (64 * TN_APPEARANCE_DIMS)
typedef struct tn_appearance {
    const float* bias;         /* [n_appearances][64] fp32 (16-byte aligned); null: no embedding */
    const uint32_t* image_id;  /* per-sample image [stride]; null: every sample takes row 0 */
    uint32_t n_appearances;    /* rows of bias; ids beyond it take row 0 */
    float* dz;                 /* backward only (may be null): dL/dz of the colour MLP's first layer,
                                  fp32 [64][stride], written for the batch's samples */
} tn_appearance_t;

/* NGP Adam settings for the hash-table range of the parameters (entries
 * [0, n_grid/2) of two parameters each), as tn_adam_step takes them: the
 * hash-grid backward can apply the step where it forms each entry's
 * gradient (tn_hashgrid_backward_planned_adam / _accumulate_adam). */
typedef struct tn_adam {
    float loss_scale;                     /* gradients are divided by it (and by 2^ceil(log2(info[0]))) */
    const uint32_t* scale_by_rays_info;   /* batch descriptor (may be null: loss_scale only) */
    float learning_rate;                  /* lr0; with d_step: lr0 / (1 + lr_decay * d_step[0]) */
    const uint32_t* d_step;
    float lr_decay, beta1, beta2, epsilon;
    float* weights;                       /* fp32 master weights (these four arrays 16-byte aligned) */
    int lp_dtype;                         /* 0: no low-precision copy, 1: fp16 copy in weights_lp */
    void* weights_lp;                     /* 8-byte aligned */
    float* first_moments;
    float* second_moments;
    uint32_t* param_steps;
} tn_adam_t;

#ifdef __cplusplus
}
#endif

#endif /* TURBONERF_TN_TYPES_H */
