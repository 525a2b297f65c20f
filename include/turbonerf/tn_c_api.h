/*
 * C ABI of the TurboNeRF MI355X kernel library (libturbonerf_kernels.so).
 *
 * Every entry point takes plain device pointers, element counts, SoA strides
 * and a hipStream_t, enqueues work on that stream and returns a tn_status
 * (0 = TN_OK). No entry point allocates device memory: workspaces are passed
 * in, with their required size reported by a matching *_bytes() query.
 * Per-ray and per-sample buffers are structure-of-arrays with component
 * stride `stride` (the reference's batch_size B), e.g. ori[i], ori[i+B],
 * ori[i+2B].
 *
 * Device-side batch descriptor `info` (uint32[4]): info[0] = rays kept in the
 * batch, info[1] = samples in the batch, info[2] = status of the selection,
 * info[3] = rays scanned (generated for the batch).
 * It is produced by tn_scan_select() and consumed by later stages, so a
 * training step needs no host round trip.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef TURBONERF_TN_C_API_H
#define TURBONERF_TN_C_API_H

#include <stddef.h>
#include <stdint.h>

#include "tn_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tn_stream_t; /* == hipStream_t */

/* ---------------------------------------------------------------- rays */

/* Seeded uniform floats in [0,1): element j = Philox-4x32-10 word
 * (offset+j)&3 of counter (offset+j)>>2 under key `seed`.
 * Replaces curandGenerateUniform (reference src/core/trainer.cu:45-50). */
int tn_fill_uniform(uint64_t n, uint64_t seed, uint64_t offset, float* out, tn_stream_t stream);

/* The next n uniforms of the same stream at a position kept on the device:
 * element j is element (*d_offset + j) of tn_fill_uniform's stream, and
 * *d_offset advances by n on the stream. A captured training step draws its
 * batch's uniforms this way, so each replay draws new ones. */
int tn_fill_uniform_dev(uint64_t n, uint64_t seed, unsigned long long* d_offset, float* out, tn_stream_t stream);

/* Training rays through random pixels, clipped to the training box.
 * Replaces initialize_training_rays_and_pixels_kernel
 * (reference src/utils/training-batch-kernels.cuh:82-187). d_n_rays: optional
 * device count (<= n_rays_max) overriding n_rays_max. */
int tn_init_training_rays(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
    uint32_t n_images, uint32_t n_pixels_per_image, int32_t img_w, int32_t img_h,
    float n_rays_per_image, const tn_bbox_t* bbox, const tn_camera_t* cameras,
    const uint8_t* images, const float* random, float* pix_rgba, uint32_t* ray_img_id,
    float* ori, float* dir, float* ray_t, float* ray_tmax, uint8_t* ray_alive,
    tn_stream_t stream);

/* Replaces deactivate_rays_with_alpha_threshold_kernel
 * (reference src/utils/training-batch-kernels.cuh:189-208). */
int tn_deactivate_rays(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
    float alpha_threshold, float selection_probability, const float* pix_rgba,
    const float* random, uint8_t* ray_alive, tn_stream_t stream);

/* ------------------------------------------------------------- marching */

/* Steps each ray takes through occupied cells. Replaces
 * march_and_count_steps_per_ray_kernel (reference
 * src/utils/training-batch-kernels.cuh:212-298). max_steps caps samples per
 * ray (0xFFFFFFFF = uncapped, the reference behaviour). */
int tn_march_count(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
    tn_bbox_t bbox, tn_occupancy_grid_t grid, float cone_angle, float dt_min, float dt_max,
    uint32_t max_steps, const float* dir, const float* ray_tmax, uint8_t* ray_alive,
    const float* ori, const float* ray_t, uint32_t* n_steps, tn_stream_t stream);

/* Exclusive scan of n_steps into ray_offset and selection of the rays that
 * form the batch; writes info[0..3] on the device. mode 0 (training): keep the
 * rays before the last one whose offset is below max_samples, the reference's
 * rule. mode 1 (rendering): keep every ray whose samples end within
 * max_samples.
 * Replaces thrust::exclusive_scan + find_last_lt_presorted + the n_samples
 * device-to-host copy (reference src/core/trainer.cu:143-183,
 * src/utils/parallel-utils.cuh:30-58). */
size_t tn_scan_workspace_bytes(uint32_t n_rays);
int tn_scan_select(int mode, uint32_t n_rays, const uint32_t* d_n_rays, const uint32_t* n_steps,
    uint32_t max_samples,
    uint32_t* ray_offset, uint32_t* info, void* workspace, size_t workspace_bytes,
    tn_stream_t stream);

/* Re-march and write compacted per-sample SoA buffers (stride sample_stride;
 * ray buffers have stride `stride`). Replaces
 * march_and_generate_network_positions_kernel (reference
 * src/utils/training-batch-kernels.cuh:339-481). */
int tn_march_generate(uint32_t n_rays_max, const uint32_t* info, uint32_t stride,
    uint32_t sample_stride, tn_bbox_t bbox, float inv_aabb, tn_occupancy_grid_t grid, float dt_min, float dt_max,
    float cone_angle, const uint32_t* ray_img_id, const float* ori, const float* dir,
    const float* ray_t, const uint32_t* ray_offset, const uint8_t* ray_alive,
    const uint32_t* n_ray_steps, uint32_t* out_img_id, float* out_t, float* out_pos,
    float* out_dir, float* out_dt, float* out_m_norm, float* out_dt_norm, tn_stream_t stream);

/* total[0] += info[1] on the device: a running count of trained samples that
 * needs no host synchronisation (no reference counterpart; the reference
 * reads n_samples back every step, src/core/trainer.cu:169-179). */
int tn_accumulate_samples(const uint32_t* info, unsigned long long* total, tn_stream_t stream);

/* tail_info = info with the sample count less `first` (0 when the batch has
 * no more samples): the descriptor of the batch's samples from `first` on,
 * for running the per-sample kernels over a batch in two parts (pointers
 * advanced by `first`, same strides); the ray count (the loss scale's) is
 * kept. No reference counterpart: the reference runs each step's kernels
 * over the whole batch (src/core/nerf-network.cu:223-317). */
int tn_batch_tail(const uint32_t* info, uint32_t first, uint32_t* tail_info, tn_stream_t stream);

/* Next step's ray count from this step's samples per ray (device-side
 * replacement of reference src/core/trainer.cu:133-140). */
int tn_estimate_rays(const uint32_t* info, uint32_t max_samples, uint32_t cap, uint32_t min_rays,
    uint32_t* n_rays_next, tn_stream_t stream);

/* ---------------------------------------------------------- compositing */

/* Fused compositing + random-background blend + smooth-L1 loss + backward.
 * Replaces fused_reconstruction_forward_backward_kernel and its memsets
 * (reference src/utils/training-kernels-fused.cuh:381-498,
 * src/core/nerf-network.cu:547-603). dtype 0: fp32 density/colour and fp32
 * colour grads; 1: fp16 in / fp16 colour grads; 2: fp32 with fp64
 * accumulation (oracle mode). density = raw network output s. Per-ray
 * buffers (bg_rgb[3], target_rgba[4], loss[4]) have stride `stride`; per-sample
 * buffers (density, color[3], dt, dL_ddensity, dL_dcolor[3]) have stride
 * `sample_stride`. dL_ddensity = dL/ds, dL_dcolor scaled by grad_scale (and by
 * 2^ceil(log2(info[0])) when scale_by_rays: the per-ray mean's 1/n_rays then
 * leaves fp16 colour gradients in range; tn_adam_step takes it back out).
 * ray_scratch (256-byte aligned, tn_composite_scratch_bytes for the ray and
 * sample capacities) holds per-ray totals and per-64-sample-chunk sums. */
size_t tn_composite_scratch_bytes(uint32_t n_rays_max, uint32_t sample_stride);
int tn_composite_train(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t stride,
    uint32_t sample_stride, float grad_scale, int scale_by_rays, float min_transmittance, const uint32_t* ray_steps,
    const uint32_t* ray_offset, const float* bg_rgb, const float* target_rgba,
    const void* density, const void* color, const float* dt, float* loss,
    float* dL_ddensity, void* dL_dcolor, void* ray_scratch, size_t scratch_bytes,
    tn_stream_t stream);

/* tn_composite_train without a workspace argument, allocating nothing: its
 * scratch is a static 32 MiB device array (one per GPU, held until the
 * process ends). No counterpart in the reference: for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). Returns TN_ERR_OUT_OF_MEMORY when
 * tn_composite_scratch_bytes(n_rays_max, sample_stride) exceeds
 * tn_composite_train_noalloc_max_bytes(). Calls that overlap on one GPU share
 * the array and race. */
size_t tn_composite_train_noalloc_max_bytes(void);
int tn_composite_train_noalloc(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t stride,
    uint32_t sample_stride, float grad_scale, int scale_by_rays, float min_transmittance,
    const uint32_t* ray_steps, const uint32_t* ray_offset, const float* bg_rgb, const float* target_rgba,
    const void* density, const void* color, const float* dt, float* loss, float* dL_ddensity,
    void* dL_dcolor, tn_stream_t stream);

/* Forward-only compositing to premultiplied rgba[4][out_stride] per ray
 * (reference density_to_ray_rgba_forward, training-kernels-fused.cuh:227-264).
 * raw_density != 0: density holds network outputs, else sigma itself. */
int tn_composite_forward(int dtype, const uint32_t* info, uint32_t stride, uint32_t max_samples,
    uint32_t out_stride, float min_transmittance, int raw_density, const uint32_t* ray_steps,
    const uint32_t* ray_offset, const void* density, const void* color, const float* dt,
    float* rgba, tn_stream_t stream);

/* mip-NeRF 360 distortion loss and dL/ds (unscaled) from raw densities,
 * normalised midpoints m_norm and steps dt_norm (reference
 * mipNeRF360_distortion_loss_forward/backward_kernel,
 * src/utils/training-kernels-fused.cuh:13-224, and its orchestration
 * src/core/nerf-network.cu:605-666; lambda is the reference's constant 0.01,
 * here a parameter). dtype 0 = fp32 density, 1 = fp16, 2 = fp64 accumulation.
 * The scratch (256-byte aligned, tn_distortion_scratch_bytes for the ray and
 * sample capacities) holds per-ray totals and per-64-sample-chunk sums. */
size_t tn_distortion_scratch_bytes(uint32_t n_rays_max, uint32_t sample_stride);
int tn_distortion(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t sample_stride,
    float lambda, const uint32_t* ray_steps, const uint32_t* ray_offset, const void* density,
    const float* m_norm, const float* dt_norm, float* ray_loss, float* dL_ds, void* scratch,
    size_t scratch_bytes, tn_stream_t stream);

/* (1/info[0]) * sum over rays and n_channels of loss -> out[0] (device).
 * Replaces the thrust::reduce of the loss buffer (reference
 * src/core/nerf-network.cu:594-602; its quirk of summing stale entries past
 * n_rays is not reproduced). */
int tn_reduce_loss(const uint32_t* info, uint32_t stride, uint32_t n_channels, const float* loss,
    float* out, tn_stream_t stream);

/* A training step's metrics in one launch: out[0] = mean over the info[0]
 * rays of the four loss channels (stride `stride`), out[1] = mean of
 * dist_loss (0 when dist_loss is null), samples_total[0] += info[1] when
 * samples_total is given, and step_counter[0] += 1 when given (the device
 * training step tn_adam_step reads for its schedule). Replaces the loss
 * reductions at the end of the reference's step (src/core/nerf-network.cu:
 * 594-602, :649-666) together with tn_accumulate_samples. */
int tn_step_metrics(const uint32_t* info, uint32_t stride, const float* loss, const float* dist_loss,
    float* out, unsigned long long* samples_total, uint32_t* step_counter, tn_stream_t stream);

/* ------------------------------------------------------- occupancy grid */

/* density *= factor (reference decay_occupancy_grid_values_kernel,
 * src/utils/occupancy-grid-kernels.cu:16-34). */
int tn_occ_decay(tn_occupancy_grid_t grid, float factor, tn_stream_t stream);

/* Jittered sample points of cells [start_idx, start_idx+n_cells) of `level`
 * (reference generate_grid_cell_network_sample_points_kernel, :38-76). */
int tn_occ_sample_points(uint32_t n_cells, uint32_t stride, uint32_t start_idx,
    tn_occupancy_grid_t grid, int level, float inv_aabb, const float* random, float* pos,
    tn_stream_t stream);

/* density = max(density, exp(min(s,12)-1)) with the warm-up/random-subset rule
 * (reference update_occupancy_with_density_kernel, :79-115). raw_dtype 0 =
 * fp32, 1 = fp16. */
int tn_occ_update_density(int raw_dtype, uint32_t n_cells, uint32_t start_idx,
    tn_occupancy_grid_t grid, int level, int sample_all_cells, const float* random,
    const void* raw_density, tn_stream_t stream);

/* bitfield byte = sum_k [density_k > threshold] << k (reference
 * update_occupancy_grid_bits_kernel, :118-141). */
int tn_occ_update_bits(tn_occupancy_grid_t grid, float threshold, tn_stream_t stream);

/* Number of set bits of a byte array -> *out (device u32). Replaces count_1s
 * (reference src/utils/bit-utils.cuh:30-72). */
size_t tn_popcount_workspace_bytes(uint64_t n_bytes);
int tn_popcount(const uint8_t* bytes, uint64_t n_bytes, uint32_t* out, void* workspace,
    size_t workspace_bytes, tn_stream_t stream);

/* ---------------------------------------------------- hash-grid encoding */

/* Level table of the multiresolution hash grid (reference HashGrid config,
 * src/core/nerf-network.cu:86-104; tiny-cuda-nn's layout is not vendored, the
 * convention is documented in hashgrid.hip). */
int tn_hashgrid_layout(int n_levels, int n_features, int log2_hashmap_size, int base_resolution,
    double per_level_scale, tn_hashgrid_t* out);

/* Encode pos[3][stride] in [0,1] into features[2L][out_stride]. table_dtype /
 * feat_dtype: 0 = fp32, 1 = fp16. Replaces the HashGrid forward inside the
 * density network (reference src/core/nerf-network.cu:461). info (optional):
 * device batch descriptor, n = min(info[1], n_max). */
int tn_hashgrid_forward(int table_dtype, int feat_dtype, uint32_t n_max, const uint32_t* info,
    uint32_t stride, uint32_t out_stride, tn_hashgrid_t grid, const float* pos,
    const void* table, void* features, tn_stream_t stream);

/* Scatter-add of feature gradients into the fp32 table gradient. Replaces the
 * HashGrid backward (reference src/core/nerf-network.cu:760). */
int tn_hashgrid_backward(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat, float* dtable,
    tn_stream_t stream);

/* The same scatter-add through a counting sort of the contributions into
 * 4096-entry bins summed in LDS (64-bit fixed point per bin), for batches of
 * up to n_max samples; same replacement as tn_hashgrid_backward (reference
 * src/core/nerf-network.cu:760). The workspace (256-byte aligned,
 * tn_hashgrid_backward_workspace_bytes, 0 = layout beyond the bin limits) is
 * scratch; without it the atomic kernel runs. */
size_t tn_hashgrid_backward_workspace_bytes(uint32_t n_max, tn_hashgrid_t grid);
int tn_hashgrid_backward_binned(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat, float* dtable,
    void* workspace, size_t workspace_bytes, tn_stream_t stream);

/* tn_hashgrid_backward_binned in two calls (same result bit for bit), for a
 * caller that has the positions before the gradients (the reference's
 * trainer does: the batch's sample positions at trainer.cu:184-216, the
 * gradients at nerf-network.cu:760). tn_hashgrid_backward_plan forms, from
 * the positions alone, the record counts and offsets of every bin into
 * `plan` (tn_hashgrid_backward_plan_bytes, 256-byte aligned); the trainer
 * runs it on its batch-generation stream. tn_hashgrid_backward_planned then
 * runs the passes that need the gradients with that plan and a work area
 * (tn_hashgrid_backward_work_bytes, 256-byte aligned). Both take the same
 * batch (n_max, info, stride, grid, pos); 0 bytes = layout beyond the bin
 * limits. */
size_t tn_hashgrid_backward_plan_bytes(uint32_t n_max, tn_hashgrid_t grid);
size_t tn_hashgrid_backward_work_bytes(uint32_t n_max, tn_hashgrid_t grid);
int tn_hashgrid_backward_plan(uint32_t n_max, const uint32_t* info, uint32_t stride, tn_hashgrid_t grid,
    const float* pos, void* plan, size_t plan_bytes, tn_stream_t stream);
int tn_hashgrid_backward_planned(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat, float* dtable,
    const void* plan, size_t plan_bytes, void* work, size_t work_bytes, tn_stream_t stream);
/* The passes of tn_hashgrid_backward_planned in parts, for a caller whose
 * feature gradients become ready by ranges of samples (the trainer's two
 * halves of a batch): tn_hashgrid_backward_scatter writes the records of
 * samples [sample_begin, sample_end) of every level (whole count blocks of
 * tn_hashgrid_backward_part_samples samples; sample_end may be n_max, and
 * the part reaching n_max must be scattered); once every part is,
 * tn_hashgrid_backward_accumulate adds the sums to dtable. Same result as
 * tn_hashgrid_backward_planned bit for bit. Same reference call site
 * (src/core/nerf-network.cu:760). */
uint32_t tn_hashgrid_backward_part_samples(uint32_t n_max, tn_hashgrid_t grid);
int tn_hashgrid_backward_scatter(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat, const void* plan,
    size_t plan_bytes, void* work, size_t work_bytes, uint32_t sample_begin, uint32_t sample_end,
    tn_stream_t stream);
int tn_hashgrid_backward_accumulate(uint32_t n_max, tn_hashgrid_t grid, float* dtable, const void* plan,
    size_t plan_bytes, void* work, size_t work_bytes, tn_stream_t stream);
/* tn_hashgrid_backward_planned / _accumulate with the optimiser step of the
 * hash table fused in: each table entry takes NGP Adam (the rule of
 * tn_adam_step: loss scale divided out, zero-gradient parameters skipped,
 * per-parameter steps, fp16 copy) where its gradient sum is formed, so the
 * table's gradient is never written and tn_adam_step then runs over the
 * network range only. Parameters, moments and steps equal those of the
 * unfused backward followed by tn_adam_step bit for bit. dtable: the fp32
 * table-gradient buffer, all zero on entry (only non-finite or oversized
 * split bins pass through it) and all zero on return. Replaces the hash
 * range of the optimiser step (reference src/core/adam-optimizer.cuh:48-96,
 * called at src/core/nerf-network.cu:770-781). */
int tn_hashgrid_backward_planned_adam(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat, float* dtable,
    const void* plan, size_t plan_bytes, void* work, size_t work_bytes, const tn_adam_t* adam,
    tn_stream_t stream);
int tn_hashgrid_backward_accumulate_adam(uint32_t n_max, tn_hashgrid_t grid, float* dtable, const void* plan,
    size_t plan_bytes, void* work, size_t work_bytes, const tn_adam_t* adam, tn_stream_t stream);


/* ------------------------------------------------------------------- MLPs */

/* Density MLP (+ SH-4 direction encoding + colour MLP when color != 0) on
 * MFMA. precision 0 = fp32, 1 = fp16 with fp32 accumulation (the first layer
 * of each MLP takes its inputs as two fp16 operands, hi + lo). params: network
 * parameters in that precision starting at element net_offset (see
 * tn_types.h); features fp32 [32][feat_stride] in both precisions; density
 * and rgb out in the network precision. Replaces the
 * FullyFusedMLP/SphericalHarmonics forward and inference (reference
 * src/core/nerf-network.cu:319-545). */
int tn_mlp_forward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t feat_stride, uint32_t out_stride, const void* params, uint32_t net_offset,
    const void* features, const float* dir, int color, void* density_out, void* rgb_out,
    const tn_appearance_t* appearance, tn_stream_t stream);

/* tn_mlp_forward for inference (rendering, the occupancy update):
 * feat_dtype 1 takes fp16 features [32][feat_stride] (precision 1 only) and
 * rounds every layer's inputs to fp16 once, as the reference's fp16 network
 * does (tcnn FullyFusedMLP inference via src/core/nerf-network.cu:431-545);
 * feat_dtype 0 is tn_mlp_forward. */
int tn_mlp_inference(int precision, int feat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t feat_stride, uint32_t out_stride, const void* params, uint32_t net_offset, const void* features,
    const float* dir, int color, void* density_out, void* rgb_out, const tn_appearance_t* appearance,
    tn_stream_t stream);

/* Backward of both MLPs with the reference's gradient plumbing (colour-MLP
 * backward, density row 0 += loss_scale*(dL_ds + dL_ds_extra), t^2 scaling for
 * t < 1, density-MLP backward; reference src/core/nerf-network.cu:668-768,
 * src/utils/training-kernels-fused.cuh:500-515). scale_by_rays: loss_scale is
 * multiplied by 2^ceil(log2(info[0])) (as tn_composite_train's grad_scale).
 * Features and the dfeat written are fp32 in both precisions; stores
 * (accumulate=0) or adds the 9408 weight gradients into grad_net (fp32). */
size_t tn_mlp_backward_workspace_bytes(uint32_t n_max);
int tn_mlp_backward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
    uint32_t feat_stride, const void* params, uint32_t net_offset, const void* features,
    const float* dir, const float* sample_t, const void* dL_drgb, uint32_t grad_stride,
    const float* dL_ds, const float* dL_ds_extra, float loss_scale, int scale_by_rays, void* dfeat,
    float* grad_net, int accumulate, void* workspace, size_t workspace_bytes,
    const tn_appearance_t* appearance, tn_stream_t stream);

/* Dynamic LDS bytes the MLP kernels launch with (for occupancy planning). */
size_t tn_mlp_lds_bytes(int precision, int backward);

/* Appearance embedding (see tn_appearance_t; `appearance` of the two calls
 * above may be null: no embedding). tn_appearance_bias forms the per-image
 * bias rows W3e emb[k] (fp32; W3e [64][16], emb [n][16]). Replaces the
 * reference's embedding forward (src/core/embedding.cuh:45-66) and its
 * concatenation into the colour input (src/core/nerf-network.cu:494-520). */
int tn_appearance_bias(uint32_t n_appearances, const float* w3e, const float* emb, float* bias,
    tn_stream_t stream);

/* Gradients of W3e and of the embedding from the colour layer's dz written
 * by tn_mlp_backward (both stored, fp32, scaled as dz is). image_id must be
 * non-decreasing over the batch's samples (a training batch's are). Replaces
 * the reference's embedding backward (src/core/embedding.cuh:68-99 and
 * nerf-network.cu:698-715), whose kernel indexes the output gradient with
 * the embedding width as its row stride and writes into the parameters. */
size_t tn_appearance_workspace_bytes(uint32_t n_appearances);
int tn_appearance_backward(uint32_t n_max, const uint32_t* info, uint32_t stride,
    const uint32_t* image_id, const float* dz, uint32_t n_appearances, const float* w3e,
    const float* emb, float* grad_w3e, float* grad_emb, void* workspace, size_t workspace_bytes,
    tn_stream_t stream);

/* ------------------------------------------------------------------ Adam */

/* NGP Adam over all parameters (reference ngp_adam_step,
 * src/core/adam-optimizer.cuh:48-96). Gradients are fp32 and are zeroed as
 * they are consumed; they are divided by loss_scale, and by
 * 2^ceil(log2(scale_by_rays_info[0])) when that batch descriptor is given.
 * The learning rate is learning_rate, or learning_rate / (1 + lr_decay *
 * *d_step) with the training step read on the device when d_step is given
 * (the reference's schedule, src/core/nerf-network.cu:256; a captured step
 * replays with the step counter advancing on the device). lp_dtype 0: no
 * low-precision copy; 1: fp16 copy. weights, gradients, moments and steps
 * must be 16-byte aligned, weights_lp 8-byte aligned. */
int tn_adam_step(uint32_t n, uint32_t n_grid, uint32_t n_net_end, float loss_scale,
    const uint32_t* scale_by_rays_info, float learning_rate, const uint32_t* d_step, float lr_decay,
    float beta1, float beta2, float epsilon, float l2_reg, float* weights,
    int lp_dtype, void* weights_lp, float* gradients, float* first_moments, float* second_moments,
    uint32_t* param_steps, tn_stream_t stream);

/* fp32 -> fp16 parameter copy (reference copy_and_cast,
 * src/utils/parallel-utils.cuh:60-82). */
int tn_cast_params(uint64_t n, const float* src, int lp_dtype, void* dst, tn_stream_t stream);

/* -------------------------------------------------------------- tuning */

/* Set / read a launch-shape knob (enum tn_tune_knob); a value outside the
 * knob's range is refused with TN_ERR_INVALID_ARGUMENT and changes nothing.
 * No reference counterpart: the reference's launch shapes are compile-time
 * constants. tn_tune_generation() counts accepted changes, so a captured HIP
 * graph (whose launch shapes are baked in) can tell that it is stale. */
int tn_tune_set(int knob, int value);
int tn_tune_get(int knob, int* value);
uint32_t tn_tune_generation(void);

/* -------------------------------------------------------------- render */

/* Row-major pixel rays for pixels [start_pixel, start_pixel + n_rays) of
 * `camera`, taken to the NeRF frame by `to_local` (optional device 3x4,
 * world -> local) and clipped to render_bbox there: origins move to the entry
 * point (ray_t = 0), ray_tmax = exit distance capped by far - near, misses are
 * not alive, pixel[i] = pixel index. Replaces
 * generate_linear_buffer_of_rays_kernel + fill_ray_buffers +
 * prepare_for_linear_raymarching_kernel (reference
 * src/rendering/ray-batch-coordinators/linear-buffer-ray-batch-coordinator.cu:11-34,
 * ray-batch-coordinator.cuh:46-84, src/utils/rendering-kernels.cu:175-260). */
int tn_render_rays(uint32_t n_rays, uint32_t start_pixel, uint32_t stride, tn_camera_t camera,
    tn_bbox_t render_bbox, const float* to_local, float* ori, float* dir, float* ray_t,
    float* ray_tmax, uint32_t* pixel, uint8_t* alive, tn_stream_t stream);

/* Interleaved RGBA image writes of the first min(info[0], n_rays) rays from
 * rgba[4][rgba_stride] (reference copy_packed_rgba_linear_buffer_kernel,
 * linear-buffer-ray-batch-coordinator.cu:54-79). */
int tn_render_write_rgba(uint32_t n_rays, const uint32_t* info, uint32_t rgba_stride,
    const uint32_t* pixel, const uint8_t* alive, const float* rgba, float* out, tn_stream_t stream);

/* Several NeRFs in one render (reference renderer.cu:344-521): world rays
 * clipped to the union of the NeRFs' boxes (prepare_for_linear_raymarching_kernel,
 * rendering-kernels.cu:175-260: origin moved to the first entry, t = 0 there,
 * t_max the last exit capped by far - near, rebased); k fixed steps per live
 * ray per pass, samples step-major (s * n_rays + i), absent steps dt < 0
 * (march_rays_and_generate_global_sample_points_kernel, :264-332); per NeRF,
 * the samples inside its box on occupied cells with their network inputs
 * (filter_and_localize_samples_for_nerf_kernel + the spatial effects +
 * assign_normalized_network_inputs_kernel, :336-478), packed to the front
 * (tn_scan_select's offsets), evaluated and added into the samples
 * (accumulate_nerf_samples_kernel, :480-528); then front-to-back compositing
 * of the summed alpha (capped at 1) and mean colour (composite_samples_kernel,
 * :531-606), ending rays at min_transmittance or their last step. */
int tn_render_rays_multi(uint32_t n_rays, uint32_t start_pixel, uint32_t stride, tn_camera_t camera,
    uint32_t n_nerfs, const tn_render_nerf_t* nerfs, float* ori, float* dir, float* ray_t,
    float* ray_tmax, uint32_t* ray_steps, uint32_t* pixel, uint8_t* alive, float* trans,
    tn_stream_t stream);
int tn_render_march_fixed(uint32_t n_rays, uint32_t k, uint32_t sample_stride, float cone_angle,
    float dt_min, float dt_max, uint32_t max_steps, const uint8_t* alive, const float* ori,
    const float* dir, uint32_t ray_stride, float* ray_t, const float* ray_tmax, uint32_t* ray_steps,
    float* s_pos, float* s_dt, tn_stream_t stream);
int tn_render_localize(uint32_t n_rays, uint32_t k, uint32_t sample_stride, tn_render_nerf_t nerf,
    tn_occupancy_grid_t grid, const float* dir, uint32_t ray_stride, const float* s_pos,
    const float* s_dt, uint32_t* valid, float* net_pos, float* net_dir, float* net_dt,
    tn_stream_t stream);
int tn_render_compact(uint32_t n_samples, uint32_t sample_stride, const uint32_t* valid,
    const uint32_t* offsets, const float* net_pos, const float* net_dir, float* c_pos, float* c_dir,
    tn_stream_t stream);
int tn_render_accumulate(int dtype, uint32_t n_samples, uint32_t sample_stride, const uint32_t* valid,
    const uint32_t* offsets, const void* density, const void* rgb, int raw_density,
    const float* net_dt, float* s_rgba, uint32_t* count, tn_stream_t stream);
int tn_render_composite_fixed(uint32_t n_rays, uint32_t k, uint32_t sample_stride,
    float min_transmittance, const float* s_dt, const float* s_rgba, const uint32_t* count,
    const uint32_t* pixel, uint8_t* alive, float* trans, float* image, tn_stream_t stream);

/* A preview: the interleaved RGBA image of a grid_w x grid_h grid of
 * pixels spread over the image, each grid pixel filling the block of image
 * pixels nearest to it (reference copy_packed_rgba_rectangular_grid_kernel,
 * src/rendering/ray-batch-coordinators/rectangular-grid-ray-batch-coordinator.cu:62-105). */
int tn_render_expand_nearest(uint32_t grid_w, uint32_t grid_h, uint32_t width, uint32_t height,
    const float* grid_rgba, float* rgba, tn_stream_t stream);

/* *out = the number of non-zero bytes in v[0, n) (reference
 * count_valued_elements, src/utils/stream-compaction.cuh:58); p[0, n) = v (in place
 * of cudaMemsetAsync, which a captured graph should not hold). */
int tn_count_nonzero(uint32_t n, const uint8_t* v, uint32_t* out, tn_stream_t stream);
int tn_fill_u32(uint32_t n, uint32_t* p, uint32_t v, tn_stream_t stream);

/* ------------------------------------------- synthetic analytic field */

/* Gaussian-blob field (blob k = {cx,cy,cz,s,p,r,g,b}): density at every cell
 * centre of every cascade (no reference counterpart: synthetic scenes). */
int tn_blob_density_grid(tn_occupancy_grid_t grid, const float* blobs, uint32_t n_blobs,
    tn_stream_t stream);

/* sigma and rgb of the blob field at normalised sample positions. */
int tn_blob_eval(const uint32_t* info, uint32_t n_max, uint32_t stride, float aabb,
    const float* blobs, uint32_t n_blobs, const float* pos, float* sigma, float* rgb,
    tn_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TURBONERF_TN_C_API_H */
