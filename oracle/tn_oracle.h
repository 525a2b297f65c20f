/*
 * TEST INFRASTRUCTURE ONLY. CPU restatement ("oracle") of the reference
 * TurboNeRF hot-path math, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. Nothing in the product
 * (turbonerf_amd/) may link, load or call this library.
 *
 * Two numeric modes:
 *   *_f32 : fp32-faithful -- same operation order as the reference, compiled
 *           with -ffp-contract=off, IEEE division/sqrt, directed-rounding
 *           division for __fdiv_ru via fesetround(FE_UPWARD).
 *   *_f64 : fp64 accumulation of the same formulas (the FP64 oracle).
 */
#ifndef TN_ORACLE_H
#define TN_ORACLE_H

#include <stdint.h>
#include "../include/turbonerf/tn_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* --- random numbers -------------------------------------------------- */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void oc_uniform(uint64_t n, uint64_t seed, uint64_t offset, float* out);

/* --- Morton ---------------------------------------------------------- */
uint32_t oc_morton3d(uint32_t x, uint32_t y, uint32_t z);
uint32_t oc_morton3d_invert(uint32_t m);

/* --- rays and marching (float, bit-exact target) -------------------- */
void oc_lens_residual_jacobian(float x, float y, float xd, float yd, const float k[5], float out[6]);
void oc_lens_undistort(float xd, float yd, const float k[5], float* xu, float* yu);
int oc_init_training_rays(uint32_t n_rays, uint32_t stride, uint32_t n_images,
    uint32_t n_pixels_per_image, int32_t img_w, int32_t img_h, float n_rays_per_image,
    const tn_bbox_t* bbox, const tn_camera_t* cameras, const uint8_t* images,
    const float* random, float* pix_rgba, uint32_t* ray_img_id, float* ori, float* dir,
    float* ray_t, float* ray_tmax, uint8_t* ray_alive);

int oc_grid_level_at(int n_levels, int resolution, float x, float y, float z, float dt);
uint32_t oc_voxel_morton_index(int resolution, int level, float x, float y, float z);
float oc_dt_to_next_voxel(int resolution, float x, float y, float z, float dx, float dy, float dz,
    float idx, float idy, float idz, float dt_min, int level);
float oc_fdiv_ru(float a, float b);

void oc_march_count(uint32_t n_rays, uint32_t stride, const tn_bbox_t* bbox,
    const uint8_t* bitfield, int n_levels, int resolution, float cone_angle, float dt_min,
    float dt_max, const float* dir, const float* ray_tmax, uint8_t* ray_alive,
    const float* ori, const float* ray_t, uint32_t* n_steps, uint32_t max_steps);

int oc_scan_select(uint32_t n_rays, const uint32_t* n_steps, uint32_t max_samples,
    uint32_t* ray_offset, uint32_t* out_n_rays, uint32_t* out_n_samples);

void oc_march_generate(uint32_t n_rays, uint32_t stride, uint32_t sample_stride, const tn_bbox_t* bbox, float inv_aabb,
    const uint8_t* bitfield, int n_levels, int resolution, float dt_min, float dt_max,
    float cone_angle, const uint32_t* ray_img_id, const float* ori, const float* dir,
    const float* ray_t, const uint32_t* ray_offset, const uint8_t* ray_alive,
    const uint32_t* n_steps, uint32_t* out_img_id, float* out_t, float* out_pos,
    float* out_dir, float* out_dt, float* out_m_norm, float* out_dt_norm);

/* --- compositing + loss + backward ----------------------------------- */
void oc_composite_train_f32(uint32_t n_rays, uint32_t stride, uint32_t sample_stride, float inv_n_rays, float grad_scale,
    const uint32_t* ray_steps, const uint32_t* ray_offset, const float* bg_rgb,
    const float* target_rgba, const float* density, const float* color, const float* dt,
    float min_transmittance, float* loss, float* dL_ddensity, float* dL_dcolor);
void oc_composite_train_f64(uint32_t n_rays, uint32_t stride, uint32_t sample_stride, double inv_n_rays, double grad_scale,
    const uint32_t* ray_steps, const uint32_t* ray_offset, const double* bg_rgb,
    const double* target_rgba, const double* density, const double* color, const double* dt,
    double min_transmittance, double* loss, double* dL_ddensity, double* dL_dcolor);

/* --- distortion loss -------------------------------------------------- */
void oc_distortion_f32(uint32_t n_rays, float lambda, const uint32_t* ray_steps,
    const uint32_t* ray_offset, const float* density, const float* m_norm, const float* dt_norm,
    float* ray_loss, float* dL_ddensity);
void oc_distortion_f64(uint32_t n_rays, double lambda, const uint32_t* ray_steps,
    const uint32_t* ray_offset, const double* density, const double* m_norm,
    const double* dt_norm, double* ray_loss, double* dL_ddensity);
double oc_distortion_bruteforce_f64(uint32_t n, double lambda, const double* density,
    const double* m_norm, const double* dt_norm);

/* --- NGP Adam --------------------------------------------------------- */
void oc_deactivate_rays(uint32_t n_rays, uint32_t stride, float alpha_threshold, float selection_probability,
                        const float* pix_rgba, const float* random, uint8_t* ray_alive);
void oc_adam_step(uint32_t n, uint32_t n_grid, uint32_t n_net_end, float loss_scale, float lr,
    float beta1, float beta2, float eps, float l2_reg, float* w_fp, float* w_lp,
    const float* grad, float* m, float* v, uint32_t* steps);

/* --- occupancy grid update -------------------------------------------- */
void oc_occ_decay(uint32_t volume, int n_levels, float factor, float* density);
void oc_occ_sample_points(uint32_t n_cells, uint32_t stride, uint32_t start_idx, int resolution,
    int level, float inv_aabb, const float* random, float* pos);
void oc_occ_update_density(uint32_t n_cells, uint32_t start_idx, uint32_t volume, int level,
    int sample_all, const float* random, const float* raw_density, const uint8_t* bitfield,
    float* density);
void oc_occ_update_bits(uint32_t volume, int n_levels, float threshold, const float* density,
    uint8_t* bitfield);
uint64_t oc_popcount_bytes(uint64_t n, const uint8_t* bytes);

/* --- hash grid, SH, MLP (fp64) ---------------------------------------- */
void oc_hashgrid_layout(int n_levels, int n_features, int log2_T, int base_res, double per_level_scale,
    tn_hashgrid_t* out);
void oc_hashgrid_fwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t stride, const float* pos,
    const double* table, double* feat);
void oc_hashgrid_bwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t stride, const float* pos,
    const double* dfeat, double* dtable);
void oc_sh4_f64(uint32_t n, uint32_t stride, const float* dir01, double* out);
void oc_nerf_fwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t stride, const float* pos,
    const float* dir01, const double* params, int color, double* density_out16, double* rgb);
/* fp32 hash grid over an fp32 table (CPU-baseline training step) */
void oc_hashgrid_fwd_f32(const tn_hashgrid_t* g, uint32_t n, uint32_t stride, const float* pos,
    const float* table, float* feat);
void oc_hashgrid_bwd_f32(const tn_hashgrid_t* g, uint32_t n, uint32_t stride, const float* pos,
    const float* dfeat, float* dtable);
/* MLP forward + backward: net = 9408 weights (W1..W5), feat [32][stride],
 * dir01 [3][stride], G = dL/drgb (loss-scaled) [3][stride] or NULL for the
 * forward only, dL_ds / extra / t [stride] (extra may be NULL); dens_out (raw
 * density output s0) [stride] and rgb_out [3][stride] may be NULL; dfeat
 * [32][stride], dW += [9408]. */
void oc_mlp_train_f32(uint32_t n, uint32_t stride, const float* net, const float* feat, const float* dir01,
    const float* G, const float* dL_ds, const float* extra, const float* t, float loss_scale, float* dens_out,
    float* rgb_out, float* dfeat, float* dW);
void oc_mlp_train_f64(uint32_t n, uint32_t stride, const double* net, const double* feat, const float* dir01,
    const double* G, const float* dL_ds, const float* extra, const float* t, double loss_scale, double* dens_out,
    double* rgb_out, double* dfeat, double* dW);

/* --- analytic blob field (synthetic scenes) --------------------------- */
void oc_blob_eval(uint32_t n, const float* xyz, uint32_t n_blobs, const float* blobs,
    float* sigma, float* rgb);
void oc_composite_analytic_f64(uint32_t n_rays, const uint32_t* ray_steps,
    const uint32_t* ray_offset, const float* sample_pos, uint32_t stride, float aabb,
    const float* sample_dt, uint32_t n_blobs, const float* blobs, double min_transmittance,
    double* rgba);

/* render rays of a pinhole camera (identity NeRF transform) */
void oc_render_rays(uint32_t n, uint32_t start, uint32_t stride, const tn_camera_t* camera,
    const tn_bbox_t* render_bbox, float* ori, float* dir, float* ray_t, float* ray_tmax, uint8_t* alive);
/* the reference's fixed-step render loop over the analytic field (fp64) */
void oc_render_fixed_step_analytic_f64(uint32_t n_rays, uint32_t stride, const float* ori, const float* dir,
    const float* ray_t, const float* ray_tmax, const tn_bbox_t* bbox, float inv_aabb, const uint8_t* bitfield,
    int n_levels, int resolution, float cone, float dt_min, float dt_max, uint32_t max_steps, uint32_t n_blobs,
    const float* blobs, double min_transmittance, double* rgba);

/* --- timing of the CPU baseline (OpenMP over rays) -------------------- */
int oc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
