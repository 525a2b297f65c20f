/*
 * TEST INFRASTRUCTURE ONLY -- the oracle. A CPU restatement of the reference
 * TurboNeRF per-ray math, used as the checker by tests/, the smoke test and
 * the cpu_baseline leg of bench.py. The product never links or calls it.
 *
 * Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off -frounding-math).
 * Each function cites the reference file:line it restates. Parity status of
 * each function (what pins it) is listed in tests/test_oracle_*.py.
 */
#include "tn_oracle.h"

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif



/* ===================================================================== */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11). Pinned by the       */
/* Random123 known-answer vectors in tests/test_oracle_rng.py.             */
/* ===================================================================== */
static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = (uint64_t)a * (uint64_t)b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
}

extern "C" void oc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(0xD2511F53u, c0, hi0, lo0);
        mulhilo32(0xCD9E8D57u, c2, hi1, lo1);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Element g of the stream (seed): block g>>2 is the Philox counter, lane g&3
 * picks the word; uniform = (word >> 8) * 2^-24 in [0, 1). This replaces the
 * reference's curandGenerateUniform (trainer.cu:45-50, Q16: seeded). */
extern "C" void oc_uniform(uint64_t n, uint64_t seed, uint64_t offset, float* out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint64_t j = 0; j < n; ++j) {
        const uint64_t g = offset + j;
        const uint64_t block = g >> 2;
        const uint32_t ctr[4] = {(uint32_t)block, (uint32_t)(block >> 32), 0u, 0u};
        uint32_t o[4];
        oc_philox4x32_10(ctr, key, o);
        out[j] = (float)(o[g & 3] >> 8) * (1.0f / 16777216.0f);
    }
}

/* ===================================================================== */
/* Morton code: textbook 3-D bit interleave (x in bit 0, y bit 1, z bit 2) */
/* as used by OccupancyGrid (reference occupancy-grid.cuh:161-175).        */
/* ===================================================================== */
extern "C" uint32_t oc_morton3d(uint32_t x, uint32_t y, uint32_t z) {
    uint32_t m = 0;
    for (int b = 0; b < 10; ++b) {
        m |= ((x >> b) & 1u) << (3 * b);
        m |= ((y >> b) & 1u) << (3 * b + 1);
        m |= ((z >> b) & 1u) << (3 * b + 2);
    }
    return m;
}

extern "C" uint32_t oc_morton3d_invert(uint32_t m) {
    uint32_t x = 0;
    for (int b = 0; b < 10; ++b) x |= ((m >> (3 * b)) & 1u) << b;
    return x;
}

/* ===================================================================== */
/* Camera rays (reference src/models/camera.cuh:94-141)                     */
/* and training ray init (src/utils/training-batch-kernels.cuh:82-187).    */
/* ===================================================================== */
struct f3 { float x, y, z; };

/* Residual and Jacobian of the radial (k1 k2 k3) + tangential (p1 p2) lens
 * model at the undistorted point (x, y), against the distorted (xd, yd)
 * (reference src/utils/camera-kernels.cuh:12-54). The reference's C++ types
 * are kept: terms written with a double literal are evaluated in double and
 * rounded to float once; the rest is float. out = fx fy fx_x fx_y fy_x fy_y. */
extern "C" void oc_lens_residual_jacobian(float x, float y, float xd, float yd, const float k[5], float out[6]) {
    const float k1 = k[0], k2 = k[1], k3 = k[2], p1 = k[3], p2 = k[4];
    const float r = x * x + y * y;
    const float d = (float)(1.0 + (double)(r * (k1 + r * (k2 + r * k3))));
    out[0] = d * x + 2.0f * p1 * x * y + p2 * (r + 2.0f * x * x) - xd;
    out[1] = d * y + 2.0f * p2 * x * y + p1 * (r + 2.0f * y * y) - yd;
    const float d_r = (float)((double)k1 + (double)r * (2.0 * (double)k2 + (double)r * (3.0 * (double)k3)));
    const float d_x = (float)(2.0 * (double)x * (double)d_r);
    const float d_y = (float)(2.0 * (double)y * (double)d_r);
    out[2] = (float)((double)(d + d_x * x) + 2.0 * (double)p1 * (double)y + 6.0 * (double)p2 * (double)x);
    out[3] = (float)((double)(d_y * x) + 2.0 * (double)p1 * (double)x + 2.0 * (double)p2 * (double)y);
    out[4] = (float)((double)(d_x * y) + 2.0 * (double)p2 * (double)y + 2.0 * (double)p1 * (double)x);
    out[5] = (float)((double)(d + d_y * y) + 2.0 * (double)p2 * (double)x + 6.0 * (double)p1 * (double)y);
}

/* Newton's method for the undistorted point: start at (xd, yd), at most 10
 * steps, stop when |det| < 1e-9 or both steps are below 1e-9
 * (reference camera-kernels.cuh:58-105, called with those constants from
 * camera.cuh:100-115). */
extern "C" void oc_lens_undistort(float xd, float yd, const float k[5], float* xu, float* yu) {
    const float eps = 1e-9f;
    float x = xd, y = yd;
    for (int it = 0; it < 10; ++it) {
        float f[6];
        oc_lens_residual_jacobian(x, y, xd, yd, k, f);
        const float det = f[3] * f[4] - f[2] * f[5];
        if (fabsf(det) < eps) break;
        const float dx = (f[0] * f[5] - f[1] * f[3]) / det;
        const float dy = (f[1] * f[2] - f[0] * f[4]) / det;
        x += dx;
        y += dy;
        if (fabsf(dx) < eps && fabsf(dy) < eps) break;
    }
    *xu = x;
    *yu = y;
}

static inline void camera_ray(const tn_camera_t& cam, float x, float y, f3& ori, f3& dir) {
    // camera.cuh:100-116: pixel coordinates of a lens-distorted camera are
    // undistorted first, in units of the focal length about the principal point
    const float* k = cam.dist_params;
    if (k[0] != 0.0f || k[1] != 0.0f || k[2] != 0.0f || k[3] != 0.0f || k[4] != 0.0f) {
        const float xd = (x - cam.principal_point[0]) / cam.focal_length[0];
        const float yd = (y - cam.principal_point[1]) / cam.focal_length[1];
        float xu, yu;
        oc_lens_undistort(xd, yd, k, &xu, &yu);
        x = xu * cam.focal_length[0] + cam.principal_point[0];
        y = yu * cam.focal_length[1] + cam.principal_point[1];
    }
    // camera.cuh:119-123
    const f3 v = {
        (x + cam.offset_px[0] + 0.5f) / cam.focal_length[0],
        (y + cam.offset_px[1] + 0.5f) / cam.focal_length[1],
        1.0f};
    // camera.cuh:126 l2_norm (tuple-math.cuh:57-66)
    const float v_len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    const float* m = cam.transform;
    // transform4f.cuh:138-145 mmul_ul3x3
    f3 gd = {m[0] * v.x + m[1] * v.y + m[2] * v.z,
             m[4] * v.x + m[5] * v.y + m[6] * v.z,
             m[8] * v.x + m[9] * v.y + m[10] * v.z};
    const float gl = sqrtf(gd.x * gd.x + gd.y * gd.y + gd.z * gd.z);
    gd = {gd.x / gl, gd.y / gl, gd.z / gl};
    // camera.cuh:135: translation + near * (v_len * global_dir)
    const f3 s = {v_len * gd.x, v_len * gd.y, v_len * gd.z};
    const f3 ns = {cam.near_plane * s.x, cam.near_plane * s.y, cam.near_plane * s.z};
    ori = {m[3] + ns.x, m[7] + ns.y, m[11] + ns.z};
    dir = gd;
}

/* BoundingBox::get_ray_t_intersections (bounding-box.cuh:27-51) */
static inline bool bbox_intersect(const tn_bbox_t& b, const f3& o, const f3& id, float& tmin, float& tmax) {
    float t1x = (b.min_x - o.x) * id.x, t2x = (b.max_x - o.x) * id.x;
    float t1y = (b.min_y - o.y) * id.y, t2y = (b.max_y - o.y) * id.y;
    float t1z = (b.min_z - o.z) * id.z, t2z = (b.max_z - o.z) * id.z;
    if (t1x > t2x) std::swap(t1x, t2x);
    if (t1y > t2y) std::swap(t1y, t2y);
    if (t1z > t2z) std::swap(t1z, t2z);
    tmin = fmaxf(t1x, fmaxf(t1y, t1z));
    tmax = fminf(t2x, fminf(t2y, t2z));
    return tmax >= tmin;
}

/* BoundingBox::contains (bounding-box.cuh:70-74) */
static inline bool bbox_contains(const tn_bbox_t& b, float x, float y, float z) {
    return x >= b.min_x && x <= b.max_x && y >= b.min_y && y <= b.max_y && z >= b.min_z && z <= b.max_z;
}

extern "C" int oc_init_training_rays(uint32_t n_rays, uint32_t B, uint32_t n_images,
    uint32_t n_pix, int32_t img_w, int32_t img_h, float n_rays_per_image,
    const tn_bbox_t* bbox, const tn_camera_t* cams, const uint8_t* images,
    const float* random, float* pix_rgba, uint32_t* ray_img_id, float* ori, float* dir,
    float* ray_t, float* ray_tmax, uint8_t* ray_alive) {
    (void)img_h;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        // training-batch-kernels.cuh:108-113 (Q5: y is not floored)
        uint32_t image_idx = (uint32_t)((float)i / n_rays_per_image);
        if (image_idx >= n_images) image_idx = n_images - 1;  // guard; unreachable for n_rays_per_image = n_rays/n_images
        const float local_pixel_idx_f = (float)n_pix * random[i];
        const float img_width_f = (float)img_w;
        const float x = fmodf(local_pixel_idx_f, img_width_f);
        const float y = local_pixel_idx_f / img_width_f;
        const tn_camera_t& cam = cams[image_idx];
        f3 o, d;
        camera_ray(cam, x, y, o, d);
        const f3 id = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        float tmin, tmax;
        if (!bbox_intersect(*bbox, o, id, tmin, tmax)) {
            ray_alive[i] = 0;
            continue;
        }
        // :144-150
        const float t_max = fminf(cam.far_plane - cam.near_plane, tmax);
        const float t = fmaxf(0.0f, tmin + 1e-5f);
        if (t_max < t) {
            ray_alive[i] = 0;
            continue;
        }
        // :159-171
        const size_t img_offset = (size_t)n_pix * image_idx;
        const size_t lp = std::min((size_t)local_pixel_idx_f, (size_t)n_pix - 1);
        const uint8_t* px = images + 4 * (img_offset + lp);
        pix_rgba[i] = (float)px[0] / 255.0f;
        pix_rgba[i + B] = (float)px[1] / 255.0f;
        pix_rgba[i + 2 * B] = (float)px[2] / 255.0f;
        pix_rgba[i + 3 * B] = (float)px[3] / 255.0f;
        ori[i] = o.x; ori[i + B] = o.y; ori[i + 2 * B] = o.z;
        dir[i] = d.x; dir[i + B] = d.y; dir[i + 2 * B] = d.z;
        ray_t[i] = t;
        ray_tmax[i] = t_max;
        ray_alive[i] = 1;
        ray_img_id[i] = image_idx;
    }
    return TN_OK;
}

/* deactivate_rays_with_alpha_threshold_kernel (reference
 * src/utils/training-batch-kernels.cuh:189-208), launched by
 * trainer.cu:85-99 with the random slice at random_float + batch_size: a ray
 * whose target alpha is below the threshold is dropped when its uniform
 * exceeds the selection probability. */
extern "C" void oc_deactivate_rays(uint32_t n_rays, uint32_t B, float alpha_threshold, float selection_probability,
                                   const float* pix_rgba, const float* random, uint8_t* ray_alive) {
    for (uint32_t i = 0; i < n_rays; ++i) {
        const float alpha = pix_rgba[3 * B + i];
        if (alpha < alpha_threshold && random[i] > selection_probability) ray_alive[i] = 0;
    }
}

/* ===================================================================== */
/* OccupancyGrid methods (reference src/core/occupancy-grid.cuh:139-282)   */
/* ===================================================================== */
extern "C" int oc_grid_level_at(int n_levels, int resolution, float x, float y, float z, float dt) {
    const float inv_res = 1.0f / (float)resolution;
    for (int k = 0; k < n_levels; ++k) {  // occupancy-grid.cuh:236-247
        const float level_size = (float)(1 << k);
        const float half = 0.5f * level_size;
        const float cell = level_size * inv_res;
        const bool covers = fabsf(x) < half && fabsf(y) < half && fabsf(z) < half;
        if (covers && cell > dt) return k;
    }
    return n_levels - 1;
}

/* occupancy-grid.cuh:150-163. Q14: the reference reads one past the grid for
 * a coordinate exactly at +extent; the index is clamped to resolution-1. */
extern "C" uint32_t oc_voxel_morton_index(int resolution, int level, float x, float y, float z) {
    const float s = 1.0f / (float)(1 << level);
    const float res_f = (float)resolution;
    const uint32_t lim = (uint32_t)resolution - 1u;
    uint32_t xu = (uint32_t)fmaxf((x * s + 0.5f) * res_f, 0.0f);
    uint32_t yu = (uint32_t)fmaxf((y * s + 0.5f) * res_f, 0.0f);
    uint32_t zu = (uint32_t)fmaxf((z * s + 0.5f) * res_f, 0.0f);
    xu = std::min(xu, lim); yu = std::min(yu, lim); zu = std::min(zu, lim);
    return oc_morton3d(xu, yu, zu);
}

/* __fdiv_ru: IEEE division rounded toward +infinity (Q4), computed here by
 * switching the FPU rounding mode. */
extern "C" float oc_fdiv_ru(float a, float b) {
    const int old = fegetround();
    fesetround(FE_UPWARD);
    volatile float va = a, vb = b;
    volatile float q = va / vb;
    fesetround(old);
    return q;
}

/* occupancy-grid.cuh:259-282 */
extern "C" float oc_dt_to_next_voxel(int resolution, float px, float py, float pz, float dx, float dy,
    float dz, float idx, float idy, float idz, float dt_min, int level) {
    const float res_f = (float)resolution;
    const float inv_res = 1.0f / res_f;
    const float level_size = (float)(1 << level);
    const float i_level_size = 1.0f / level_size;
    const float x = (px * i_level_size + 0.5f) * res_f;
    const float y = (py * i_level_size + 0.5f) * res_f;
    const float z = (pz * i_level_size + 0.5f) * res_f;
    const float k = inv_res * level_size;
    const float tx = k * ((floorf(0.5f * copysignf(1.0f, dx) + x + 0.5f) - x) * idx);
    const float ty = k * ((floorf(0.5f * copysignf(1.0f, dy) + y + 0.5f) - y) * idy);
    const float tz = k * ((floorf(0.5f * copysignf(1.0f, dz) + z + 0.5f) - z) * idz);
    const float t_target = fmaxf(dt_min, fminf(fminf(tx, ty), tz));
    return oc_fdiv_ru(t_target, dt_min) * dt_min;
}

/* raymarching-utils.cuh:7-14 */
static inline float get_dt(float t, float cone, float dt_min, float dt_max) {
    return std::min(std::max(t * cone, dt_min), dt_max);
}

static inline bool grid_occupied(const uint8_t* bits, int res, int level, float x, float y, float z) {
    return (bits[oc_voxel_morton_index(res, level, x, y, z)] >> level) & 1u;
}

/* training-batch-kernels.cuh:212-298. max_steps caps samples per ray
 * (0xFFFFFFFF = no cap, the reference behaviour). */
extern "C" void oc_march_count(uint32_t n_rays, uint32_t B, const tn_bbox_t* bbox,
    const uint8_t* bits, int n_levels, int res, float cone, float dt_min, float dt_max,
    const float* dir, const float* ray_tmax, uint8_t* ray_alive, const float* ori,
    const float* ray_t, uint32_t* n_steps, uint32_t max_steps) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (!ray_alive[i]) { n_steps[i] = 0; continue; }
        const float ox = ori[i], oy = ori[i + B], oz = ori[i + 2 * B];
        const float dx = dir[i], dy = dir[i + B], dz = dir[i + 2 * B];
        const float idx = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
        uint32_t n = 0;
        float t = ray_t[i];
        const float t_max = ray_tmax[i];
        while (t < t_max && n < max_steps) {
            const float x = ox + t * dx, y = oy + t * dy, z = oz + t * dz;
            if (!bbox_contains(*bbox, x, y, z)) break;
            const float dt = get_dt(t, cone, dt_min, dt_max);
            const int lvl = oc_grid_level_at(n_levels, res, x, y, z, dt);
            if (grid_occupied(bits, res, lvl, x, y, z)) {
                t += dt;
                ++n;
            } else {
                t += oc_dt_to_next_voxel(res, x, y, z, dx, dy, dz, idx, idy, idz, dt_min, lvl);
            }
        }
        if (n == 0) ray_alive[i] = 0;
        n_steps[i] = n;
    }
}

/* trainer.cu:143-183 + parallel-utils.cuh:30-58: exclusive scan, then keep the
 * largest prefix of rays whose offset stays below max_samples (Q8: the scan
 * covers exactly the n_rays generated this step). */
extern "C" int oc_scan_select(uint32_t n_rays, const uint32_t* n_steps, uint32_t max_samples,
    uint32_t* ray_offset, uint32_t* out_n_rays, uint32_t* out_n_samples) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n_rays; ++i) {
        ray_offset[i] = (uint32_t)std::min<uint64_t>(acc, 0xFFFFFFFFull);
        acc += n_steps[i];
    }
    int last = -1;
    for (int i = (int)n_rays - 1; i >= 0; --i) {
        if (ray_offset[i] < max_samples) { last = i; break; }
    }
    const int n_ray_max_idx = last - 1;
    if (n_ray_max_idx < 0) { *out_n_rays = 0; *out_n_samples = 0; return TN_ERR_NO_RAYS; }
    *out_n_rays = (uint32_t)(n_ray_max_idx + 1);
    *out_n_samples = ray_offset[*out_n_rays];
    if (*out_n_samples < 1) return TN_ERR_NO_SAMPLES;
    return TN_OK;
}

/* training-batch-kernels.cuh:305-331 */
static inline void assign_normalized(uint32_t SB, uint32_t sid, float x, float y, float z, float dx, float dy,
    float dz, float dt, float inv_aabb, float* pos, float* dir, float* odt) {
    odt[sid] = inv_aabb * dt;
    pos[sid] = std::min(std::max(x * inv_aabb + 0.5f, 0.0f), 1.0f);
    pos[sid + SB] = std::min(std::max(y * inv_aabb + 0.5f, 0.0f), 1.0f);
    pos[sid + 2 * SB] = std::min(std::max(z * inv_aabb + 0.5f, 0.0f), 1.0f);
    dir[sid] = 0.5f * dx + 0.5f;
    dir[sid + SB] = 0.5f * dy + 0.5f;
    dir[sid + 2 * SB] = 0.5f * dz + 0.5f;
}

/* training-batch-kernels.cuh:339-481 */
extern "C" void oc_march_generate(uint32_t n_rays, uint32_t B, uint32_t SB, const tn_bbox_t* bbox, float inv_aabb,
    const uint8_t* bits, int n_levels, int res, float dt_min, float dt_max, float cone,
    const uint32_t* ray_img_id, const float* ori, const float* dir, const float* ray_t,
    const uint32_t* ray_offset, const uint8_t* ray_alive, const uint32_t* n_ray_steps,
    uint32_t* out_img_id, float* out_t, float* out_pos, float* out_dir, float* out_dt,
    float* out_m_norm, float* out_dt_norm) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        if (!ray_alive[i]) continue;
        const uint32_t n_steps = n_ray_steps[i];
        const uint32_t off = ray_offset[i];
        const uint32_t img = ray_img_id[i];
        const float ox = ori[i], oy = ori[i + B], oz = ori[i + 2 * B];
        const float dx = dir[i], dy = dir[i + B], dz = dir[i + 2 * B];
        const float idx = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
        float t = ray_t[i];
        const float t_min = t;
        float t_max = t;
        uint32_t n = 0;
        while (n < n_steps) {
            const float x = ox + t * dx, y = oy + t * dy, z = oz + t * dz;
            const float dt = get_dt(t, cone, dt_min, dt_max);
            const uint32_t sid = off + n;
            if (!bbox_contains(*bbox, x, y, z)) {  // :422-439 (Q20: unreachable)
                out_t[sid] = t; out_m_norm[sid] = t + 0.5f * dt; out_dt_norm[sid] = dt; out_img_id[sid] = img;
                assign_normalized(SB, sid, x, y, z, dx, dy, dz, dt, inv_aabb, out_pos, out_dir, out_dt);
                ++n;
                continue;
            }
            const int lvl = oc_grid_level_at(n_levels, res, x, y, z, dt);
            if (grid_occupied(bits, res, lvl, x, y, z)) {
                out_t[sid] = t; out_m_norm[sid] = t + 0.5f * dt; out_dt_norm[sid] = dt; out_img_id[sid] = img;
                t += dt;
                t_max = t;
                assign_normalized(SB, sid, x, y, z, dx, dy, dz, dt, inv_aabb, out_pos, out_dir, out_dt);
                ++n;
            } else {
                t += oc_dt_to_next_voxel(res, x, y, z, dx, dy, dz, idx, idy, idz, dt_min, lvl);
            }
        }
        const float t_span = std::max(t_max - t_min, 1e-6f);  // :475-480
        for (uint32_t j = 0; j < n; ++j) {
            out_m_norm[off + j] = std::min(std::max((out_m_norm[off + j] - t_min) / t_span, 0.0f), 1.0f);
            out_dt_norm[off + j] /= t_span;
        }
    }
}

/* ===================================================================== */
/* Compositing + smooth-L1 + backward                                     */
/* (reference training-kernels-fused.cuh:227-498, common-network-kernels  */
/*  .cuh:7-17). Q2 (bg composite of premultiplied R), Q3 (1/n_rays), Q6    */
/* (dsigma/ds = sigma) reproduced.                                         */
/* ===================================================================== */
template <typename R> static inline R r_exp(R x) { return std::exp(x); }
template <> inline float r_exp<float>(float x) { return expf(x); }

template <typename R>
static void composite_train(uint32_t n_rays, uint32_t B, uint32_t SB, R inv_n_rays, R grad_scale,
    const uint32_t* steps, const uint32_t* offs, const R* bg, const R* target, const R* density,
    const R* color, const R* dtb, R min_T, R* loss, R* dL_ddensity, R* dL_dcolor) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const uint32_t n = steps[i];
        const uint32_t o = offs[i];
        const R* sr = color + o; const R* sg = color + o + SB; const R* sb = color + o + 2 * (size_t)SB;
        const R* sd = density + o; const R* sdt = dtb + o;
        // density_to_ray_rgba_forward (:227-264)
        R r = 0, g = 0, b = 0, a = 0, T = 1;
        for (uint32_t k = 0; k < n; ++k) {
            const R sigma = r_exp<R>(std::min(sd[k], (R)12) - (R)1);
            const R alpha = (R)1 - r_exp<R>(-sigma * sdt[k]);
            const R w = T * alpha;
            r += w * sr[k]; g += w * sg[k]; b += w * sb[k]; a += w;
            T *= ((R)1 - alpha);
            if (T < min_T) break;
        }
        // :434-456 random-background compositing
        const R gr = target[i], gg = target[i + B], gb = target[i + 2 * B], ga = target[i + 3 * B];
        const R rr = bg[i], rg = bg[i + B], rb = bg[i + 2 * B];
        const R gac = (R)1 - ga;
        const R grc = gr * ga + rr * gac, ggc = gg * ga + rg * gac, gbc = gb * ga + rb * gac;
        const R rac = (R)1 - a;
        const R rrc = r * a + rr * rac, rgc = g * a + rg * rac, rbc = b * a + rb * rac;
        const R dr = rrc - grc, dg = rgc - ggc, db = rbc - gbc, da = rac - gac;
        auto sl1 = [](R x) { const R ax = std::fabs(x); return ax < (R)1 ? (R)0.5 * x * x : ax - (R)0.5; };
        auto sl1b = [](R x) { return std::fabs(x) < (R)1 ? x : std::copysign((R)1, x); };
        loss[i] = sl1(dr); loss[i + B] = sl1(dg); loss[i + 2 * B] = sl1(db); loss[i + 3 * B] = sl1(da);
        const R dLr = inv_n_rays * sl1b(dr), dLg = inv_n_rays * sl1b(dg), dLb = inv_n_rays * sl1b(db);
        // density_to_ray_rgba_backward (:269-341)
        const R ddr = r - rr, ddg = g - rg, ddb = b - rb;
        R cr = r, cg = g, cb = b, ca = a;
        T = 1;
        for (uint32_t k = 0; k < n; ++k) {
            const R cR = sr[k], cG = sg[k], cB = sb[k];
            const R dt = sdt[k];
            const R sigma = r_exp<R>(std::min(sd[k], (R)12) - (R)1);
            const R alpha = (R)1 - r_exp<R>(-sigma * dt);
            const R w = T * alpha;
            const R kk = T - w;
            cr -= w * cR; cg -= w * cG; cb -= w * cB; ca -= w;
            R dRr = dt * (kk * cR - cr), dRg = dt * (kk * cG - cg), dRb = dt * (kk * cB - cb);
            const R dRa = dt * (kk * (R)1 - ca);
            dRr = dRr * a + ddr * dRa; dRg = dRg * a + ddg * dRa; dRb = dRb * a + ddb * dRa;
            const R dL_dsigma = dLr * dRr + dLg * dRg + dLb * dRb;
            dL_ddensity[o + k] = dL_dsigma * sigma;
            const R dRdc = w * a;
            dL_dcolor[o + k] = grad_scale * dLr * dRdc;
            dL_dcolor[o + k + SB] = grad_scale * dLg * dRdc;
            dL_dcolor[o + k + 2 * (size_t)SB] = grad_scale * dLb * dRdc;
            T *= (R)1 - alpha;
            if (T < min_T) break;
        }
    }
}

extern "C" void oc_composite_train_f32(uint32_t n_rays, uint32_t B, uint32_t SB, float inv_n_rays, float grad_scale,
    const uint32_t* steps, const uint32_t* offs, const float* bg, const float* target,
    const float* density, const float* color, const float* dt, float min_T, float* loss,
    float* dL_ddensity, float* dL_dcolor) {
    composite_train<float>(n_rays, B, SB, inv_n_rays, grad_scale, steps, offs, bg, target, density, color,
        dt, min_T, loss, dL_ddensity, dL_dcolor);
}

extern "C" void oc_composite_train_f64(uint32_t n_rays, uint32_t B, uint32_t SB, double inv_n_rays, double grad_scale,
    const uint32_t* steps, const uint32_t* offs, const double* bg, const double* target,
    const double* density, const double* color, const double* dt, double min_T, double* loss,
    double* dL_ddensity, double* dL_dcolor) {
    composite_train<double>(n_rays, B, SB, inv_n_rays, grad_scale, steps, offs, bg, target, density, color,
        dt, min_T, loss, dL_ddensity, dL_dcolor);
}

/* ===================================================================== */
/* mip-NeRF 360 distortion loss, O(N) prefix-sum form                      */
/* (reference training-kernels-fused.cuh:13-224). lambda is a parameter    */
/* (reference constant 0.01, nerf-constants.cuh:17). No early stop (Q17).  */
/* ===================================================================== */
template <typename R>
static void distortion(uint32_t n_rays, R lambda, const uint32_t* steps, const uint32_t* offs,
    const R* density, const R* mn, const R* dtn, R* ray_loss, R* dL_ddensity) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const uint32_t n = steps[i];
        if (n == 0) { ray_loss[i] = 0; continue; }
        const uint32_t o = offs[i];
        const R* d = density + o; const R* dt = dtn + o; const R* m = mn + o;
        std::vector<R> w_cs(n), wm_cs(n);
        R cw = 0, cwm = 0, cdtw2 = 0, cwmw = 0, cwwm = 0, T = 1, pw = 0, pwm = 0, b0 = 0, b1 = 0, uni = 0;
        for (uint32_t k = 0; k < n; ++k) {  // forward :81-111
            const R sigma = r_exp<R>(std::min(d[k], (R)12) - (R)1);
            const R alpha = (R)1 - r_exp<R>(-sigma * dt[k]);
            const R w = alpha * T;
            const R wm = w * m[k];
            const R a0 = wm * pw, a1 = w * pwm, dtw2 = dt[k] * w * w;
            cw += w; cwm += wm; cdtw2 += dtw2; cwmw += a0; cwwm += a1;
            w_cs[k] = cw; wm_cs[k] = cwm;
            b0 += a0; b1 += a1; uni += dtw2;
            pw = cw; pwm = cwm;
            T *= ((R)1 - alpha);
        }
        const R kf = lambda / (R)n;
        ray_loss[i] = kf * ((R)(1.0 / 3.0) * uni + (R)2 * (b0 - b1));
        T = 1; pw = 0; pwm = 0;
        for (uint32_t k = 0; k < n; ++k) {  // backward :188-223
            const R sigma = r_exp<R>(std::min(d[k], (R)12) - (R)1);
            const R alpha = (R)1 - r_exp<R>(-sigma * dt[k]);
            const R Ti = T;
            const R w = alpha * Ti;
            const R mi = m[k];
            const R wm = w * mi;
            const R dalpha = dt[k] * ((R)1 - alpha);
            const R dw = Ti * dalpha;
            const R dwm = mi * dw;
            const R tw = Ti - w;
            cw -= w; cwm -= wm; cdtw2 -= dt[k] * w * w; cwmw -= wm * pw; cwwm -= w * pwm;
            const R db0 = dwm * pw + dt[k] * (cwm * (w_cs[k] + tw) - (R)2 * cwmw);
            const R db1 = dw * pwm + dt[k] * (cw * (mi * tw + wm_cs[k]) - (R)2 * cwwm);
            const R duni = (R)2 * dt[k] * (dt[k] * w * tw - cdtw2);
            const R dsig = kf * ((R)2 * (db0 - db1) + ((R)1 / (R)3) * duni);
            dL_ddensity[o + k] = dsig * sigma;
            pw = w_cs[k]; pwm = wm_cs[k];
            T *= (R)1 - alpha;
        }
    }
}

extern "C" void oc_distortion_f32(uint32_t n_rays, float lambda, const uint32_t* steps,
    const uint32_t* offs, const float* density, const float* m, const float* dt, float* ray_loss,
    float* dL) {
    distortion<float>(n_rays, lambda, steps, offs, density, m, dt, ray_loss, dL);
}

extern "C" void oc_distortion_f64(uint32_t n_rays, double lambda, const uint32_t* steps,
    const uint32_t* offs, const double* density, const double* m, const double* dt,
    double* ray_loss, double* dL) {
    distortion<double>(n_rays, lambda, steps, offs, density, m, dt, ray_loss, dL);
}

/* mip-NeRF 360 eq. (15) as the O(N^2) double sum; used to pin the O(N) form. */
extern "C" double oc_distortion_bruteforce_f64(uint32_t n, double lambda, const double* d,
    const double* m, const double* dt) {
    std::vector<double> w(n);
    double T = 1;
    for (uint32_t k = 0; k < n; ++k) {
        const double sigma = std::exp(std::min(d[k], 12.0) - 1.0);
        const double alpha = 1.0 - std::exp(-sigma * dt[k]);
        w[k] = T * alpha;
        T *= 1.0 - alpha;
    }
    double bi = 0, uni = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uni += dt[i] * w[i] * w[i];
        for (uint32_t j = 0; j < n; ++j) bi += w[i] * w[j] * std::fabs(m[i] - m[j]);
    }
    return lambda / (double)n * (uni / 3.0 + bi);
}

/* ===================================================================== */
/* NGP Adam (reference src/core/adam-optimizer.cuh:48-96)                  */
/* ===================================================================== */
extern "C" void oc_adam_step(uint32_t n, uint32_t n_grid, uint32_t n_net_end, float loss_scale, float lr0,
    float beta1, float beta2, float eps, float l2_reg, float* w_fp, float* w_lp, const float* grad,
    float* m1, float* m2, uint32_t* steps) {
    for (uint32_t i = 0; i < n; ++i) {
        float g = grad[i] / loss_scale;
        if (i < n_grid && g == 0) continue;
        const float wf = w_fp[i];
        if (i >= n_grid && i < n_net_end) g += l2_reg * wf;
        const float g2 = g * g;
        const float fm = m1[i] = beta1 * m1[i] + (1 - beta1) * g;
        const float sm = m2[i] = beta2 * m2[i] + (1 - beta2) * g2;
        const uint32_t st = ++steps[i];
        float lr = lr0;
        lr *= sqrtf(1 - powf(beta2, (float)st)) / (1 - powf(beta1, (float)st));
        const float elr = lr / (sqrtf(sm) + eps);
        const float nw = wf - elr * fm;
        w_fp[i] = nw;
        w_lp[i] = nw;
    }
}

/* ===================================================================== */
/* Occupancy grid update (reference src/utils/occupancy-grid-kernels.cu)   */
/* ===================================================================== */
extern "C" void oc_occ_decay(uint32_t volume, int n_levels, float factor, float* density) {
    for (int l = 0; l < n_levels; ++l)  // :16-34
        for (uint32_t i = 0; i < volume; ++i) density[(size_t)l * volume + i] *= factor;
}

/* :38-76 with get_voxel_xyz_from_morton_index (occupancy-grid.cuh:167-182).
 * Q13: positions are not clamped. */
extern "C" void oc_occ_sample_points(uint32_t n_cells, uint32_t B, uint32_t start_idx, int res, int level,
    float inv_aabb, const float* random, float* pos) {
    const float inv_res = 1.0f / (float)res;
    const int half_res = res / 2;
    for (uint32_t i = 0; i < n_cells; ++i) {
        const uint32_t idx = i + start_idx;
        const float level_size = (float)(1 << level);
        const float voxel = level_size * inv_res;
        const int ix = (int)oc_morton3d_invert(idx >> 0);
        const int iy = (int)oc_morton3d_invert(idx >> 1);
        const int iz = (int)oc_morton3d_invert(idx >> 2);
        const float vx = (float)(ix - half_res) * voxel;
        const float vy = (float)(iy - half_res) * voxel;
        const float vz = (float)(iz - half_res) * voxel;
        const float x = vx + random[i] * voxel;
        const float y = vy + random[i + B] * voxel;
        const float z = vz + random[i + 2 * B] * voxel;
        pos[i] = (x * inv_aabb) + 0.5f;
        pos[i + B] = (y * inv_aabb) + 0.5f;
        pos[i + 2 * B] = (z * inv_aabb) + 0.5f;
    }
}

/* :79-115 (Q12: decay + max, random subset once step >= 256) */
extern "C" void oc_occ_update_density(uint32_t n_cells, uint32_t start_idx, uint32_t volume, int level,
    int sample_all, const float* random, const float* raw, const uint8_t* bits, float* density) {
    for (uint32_t i = 0; i < n_cells; ++i) {
        const uint32_t idx = i + start_idx;
        if (!sample_all) {
            const float r = random[i];
            if (r > 0.5f) continue;
            if (!(r < 0.25f)) {
                if (!((bits[idx] >> level) & 1u)) continue;
            }
        }
        const float sigma = expf(std::min(raw[i], 12.0f) - 1.0f);
        float& dv = density[(size_t)level * volume + idx];
        dv = std::max(sigma, dv);
    }
}

/* :118-141 */
extern "C" void oc_occ_update_bits(uint32_t volume, int n_levels, float thr, const float* density,
    uint8_t* bits) {
    for (uint32_t i = 0; i < volume; ++i) {
        uint8_t b = 0;
        for (int l = 0; l < n_levels; ++l) b |= (uint8_t)((density[(size_t)l * volume + i] > thr ? 1 : 0) << l);
        bits[i] = b;
    }
}

/* bit-utils.cuh:30-72 (count_1s) */
extern "C" uint64_t oc_popcount_bytes(uint64_t n, const uint8_t* bytes) {
    uint64_t c = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t v = bytes[i];
        while (v) { c += v & 1u; v >>= 1; }
    }
    return c;
}

/* ===================================================================== */
/* Multiresolution hash grid (Instant-NGP, Mueller et al. 2022 sec. 3;    */
/* configuration at reference nerf-network.cu:86-104). tiny-cuda-nn is    */
/* absent, so the per-level layout below is this project's documented     */
/* convention (parity unpinned against tcnn; pinned by properties).        */
/* ===================================================================== */
extern "C" void oc_hashgrid_layout(int L, int F, int log2_T, int base_res, double b, tn_hashgrid_t* g) {
    std::memset(g, 0, sizeof(*g));
    g->n_levels = L; g->n_features = F; g->log2_hashmap_size = log2_T; g->base_resolution = base_res;
    g->per_level_scale = (float)b;
    const uint64_t T = 1ull << log2_T;
    uint64_t off = 0;
    for (int l = 0; l < L; ++l) {
        const double s = std::exp2((double)l * std::log2(b)) * (double)base_res - 1.0;
        const float sf = (float)s;
        const uint32_t res = (uint32_t)std::ceil((double)sf) + 1u;
        const uint64_t vol = (uint64_t)res * res * res;
        uint64_t size = (vol + 7) / 8 * 8;
        if (size > T) size = T;
        g->scale[l] = sf;
        g->resolution[l] = res;
        g->level_size[l] = (uint32_t)size;
        g->level_offset[l] = (uint32_t)off;
        g->level_dense[l] = vol <= size ? 1u : 0u;
        off += size;
    }
    g->n_params = (uint32_t)(off * (uint64_t)F);
}

static inline uint32_t grid_index(const tn_hashgrid_t* g, int l, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t idx;
    if (g->level_dense[l]) {
        const uint32_t r = g->resolution[l];
        idx = x + y * r + z * r * r;
    } else {
        idx = (x * 1u) ^ (y * 2654435761u) ^ (z * 805459861u);
    }
    return idx % g->level_size[l];
}

/* cell index and fraction are formed in fp32 (x*scale + 0.5f, floorf) so the
 * oracle selects exactly the cells the device kernel selects; weights and
 * sums are fp64. */
template <typename F>
static void hash_visit(const tn_hashgrid_t* g, int l, float px, float py, float pz, F&& f) {
    const float s = g->scale[l];
    const float x = px * s + 0.5f, y = py * s + 0.5f, z = pz * s + 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const double fx = (double)(x - fx0), fy = (double)(y - fy0), fz = (double)(z - fz0);
    const uint32_t gx = (uint32_t)(int32_t)fx0, gy = (uint32_t)(int32_t)fy0, gz = (uint32_t)(int32_t)fz0;
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const double w = (cx ? fx : 1 - fx) * (cy ? fy : 1 - fy) * (cz ? fz : 1 - fz);
        f(g->level_offset[l] + grid_index(g, l, gx + cx, gy + cy, gz + cz), w);
    }
}

extern "C" void oc_hashgrid_fwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t B, const float* pos,
    const double* table, double* feat) {
    const int F = g->n_features;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        for (int l = 0; l < g->n_levels; ++l) {
            double acc[8] = {0};
            hash_visit(g, l, pos[i], pos[i + B], pos[i + 2 * B], [&](uint32_t e, double w) {
                for (int f = 0; f < F; ++f) acc[f] += w * table[(size_t)e * F + f];
            });
            for (int f = 0; f < F; ++f) feat[(size_t)(l * F + f) * B + i] = acc[f];
        }
    }
}

extern "C" void oc_hashgrid_bwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t B, const float* pos,
    const double* dfeat, double* dtable) {
    const int F = g->n_features;
    for (uint32_t i = 0; i < n; ++i) {
        for (int l = 0; l < g->n_levels; ++l) {
            hash_visit(g, l, pos[i], pos[i + B], pos[i + 2 * B], [&](uint32_t e, double w) {
                for (int f = 0; f < F; ++f) dtable[(size_t)e * F + f] += w * dfeat[(size_t)(l * F + f) * B + i];
            });
        }
    }
}

/* Real spherical harmonics up to degree 4 (16 coefficients), standard
 * orthonormal basis with Condon-Shortley phase (Sloan, "Efficient Spherical
 * Harmonic Evaluation", JCGT 2013). Input is 0.5*d + 0.5 as the reference
 * feeds it (training-batch-kernels.cuh:328-330). */
extern "C" void oc_sh4_f64(uint32_t n, uint32_t B, const float* dir01, double* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const double x = 2.0 * dir01[i] - 1.0, y = 2.0 * dir01[i + B] - 1.0, z = 2.0 * dir01[i + 2 * B] - 1.0;
        const double xx = x * x, yy = y * y, zz = z * z;
        double* o = out + i;
        o[0 * B] = 0.28209479177387814;
        o[1 * B] = -0.48860251190291987 * y;
        o[2 * B] = 0.48860251190291987 * z;
        o[3 * B] = -0.48860251190291987 * x;
        o[4 * B] = 1.0925484305920792 * x * y;
        o[5 * B] = -1.0925484305920792 * y * z;
        o[6 * B] = 0.94617469575755997 * zz - 0.31539156525251999;
        o[7 * B] = -1.0925484305920792 * x * z;
        o[8 * B] = 0.54627421529603959 * (xx - yy);
        o[9 * B] = 0.59004358992664352 * y * (-3.0 * xx + yy);
        o[10 * B] = 2.8906114426405538 * x * y * z;
        o[11 * B] = 0.45704579946446572 * y * (1.0 - 5.0 * zz);
        o[12 * B] = 0.3731763325901154 * z * (5.0 * zz - 3.0);
        o[13 * B] = 0.45704579946446572 * x * (1.0 - 5.0 * zz);
        o[14 * B] = 1.4453057213202769 * z * (xx - yy);
        o[15 * B] = 0.59004358992664352 * x * (-xx + 3.0 * yy);
    }
}

/* Density MLP 32->64(ReLU)->16 and colour MLP 32->64->64->3 (sigmoid), bias
 * free, row-major [out][in] weights after the hash table (parameter layout in
 * tn_types.h; network shapes from reference nerf-network.cu:61-73,106-121). */
extern "C" void oc_nerf_fwd_f64(const tn_hashgrid_t* g, uint32_t n, uint32_t B, const float* pos,
    const float* dir01, const double* params, int color, double* d16, double* rgb) {
    std::vector<double> feat((size_t)32 * n), sh((size_t)16 * n);
    oc_hashgrid_fwd_f64(g, n, n, pos, params, feat.data());
    (void)B;
    const double* W1 = params + g->n_params;
    const double* W2 = W1 + 64 * 32;
    const double* W3 = W2 + 16 * 64;
    const double* W4 = W3 + 64 * 32;
    const double* W5 = W4 + 64 * 64;
    if (color) oc_sh4_f64(n, B, dir01, sh.data());
    for (uint32_t i = 0; i < n; ++i) {
        double x[32], h[64], h2[64], o16[16];
        for (int k = 0; k < 32; ++k) x[k] = feat[(size_t)k * n + i];
        for (int r = 0; r < 64; ++r) { double a = 0; for (int k = 0; k < 32; ++k) a += W1[r * 32 + k] * x[k]; h[r] = a > 0 ? a : 0; }
        for (int r = 0; r < 16; ++r) { double a = 0; for (int k = 0; k < 64; ++k) a += W2[r * 64 + k] * h[k]; o16[r] = a; }
        for (int r = 0; r < 16; ++r) d16[(size_t)r * B + i] = o16[r];
        if (!color) continue;
        double ci[32];
        for (int k = 0; k < 16; ++k) ci[k] = o16[k];
        for (int k = 0; k < 16; ++k) ci[16 + k] = sh[(size_t)k * B + i];
        for (int r = 0; r < 64; ++r) { double a = 0; for (int k = 0; k < 32; ++k) a += W3[r * 32 + k] * ci[k]; h[r] = a > 0 ? a : 0; }
        for (int r = 0; r < 64; ++r) { double a = 0; for (int k = 0; k < 64; ++k) a += W4[r * 64 + k] * h[k]; h2[r] = a > 0 ? a : 0; }
        for (int r = 0; r < 3; ++r) { double a = 0; for (int k = 0; k < 64; ++k) a += W5[r * 64 + k] * h2[k]; rgb[(size_t)r * B + i] = 1.0 / (1.0 + std::exp(-a)); }
    }
}

/* fp32 hash-grid encoding over an fp32 table (the CPU baseline's training
 * step); backward scatter-adds with atomics, as every parallel CPU port must. */
extern "C" void oc_hashgrid_fwd_f32(const tn_hashgrid_t* g, uint32_t n, uint32_t B, const float* pos,
    const float* table, float* feat) {
    const int F = g->n_features;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        for (int l = 0; l < g->n_levels; ++l) {
            float acc[8] = {0};
            hash_visit(g, l, pos[i], pos[i + B], pos[i + 2 * B], [&](uint32_t e, double w) {
                for (int f = 0; f < F; ++f) acc[f] += (float)w * table[(size_t)e * F + f];
            });
            for (int f = 0; f < F; ++f) feat[(size_t)(l * F + f) * B + i] = acc[f];
        }
    }
}

extern "C" void oc_hashgrid_bwd_f32(const tn_hashgrid_t* g, uint32_t n, uint32_t B, const float* pos,
    const float* dfeat, float* dtable) {
    const int F = g->n_features;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        for (int l = 0; l < g->n_levels; ++l) {
            hash_visit(g, l, pos[i], pos[i + B], pos[i + 2 * B], [&](uint32_t e, double w) {
                for (int f = 0; f < F; ++f) {
                    const float v = (float)w * dfeat[(size_t)(l * F + f) * B + i];
#pragma omp atomic
                    dtable[(size_t)e * F + f] += v;
                }
            });
        }
    }
}

/* Density + colour MLP forward and backward for one batch of samples
 * (reference nerf-network.cu:431-545 forward, :668-768 backward):
 *   dL/do = G_rgb * rgb * (1 - rgb) through the sigmoid;
 *   dS = W3^T dH2 restricted to the 16 density outputs;
 *   dS[0] += loss_scale * (dL/ds + extra); then dS[0] *= t^2 where t < 1;
 *   weight gradients accumulate into dW (+=), feature gradients go to dfeat.
 * R is the arithmetic type. OpenMP over samples with one dW accumulator per
 * thread, summed in thread order. */
template <typename R>
static void mlp_train(uint32_t n, uint32_t B, const R* net, const R* feat, const float* dir01, const R* G,
    const float* dL_ds, const float* extra, const float* t, R loss_scale, R* dens_out, R* rgb_out, R* dfeat, R* dW) {
    const R* W1 = net;
    const R* W2 = W1 + 64 * 32;
    const R* W3 = W2 + 16 * 64;
    const R* W4 = W3 + 64 * 32;
    const R* W5 = W4 + 64 * 64;
    const int NP = 9408;
    const int nt = omp_get_max_threads();
    std::vector<R> part((size_t)nt * NP, (R)0);
#pragma omp parallel
    {
        R* gw = part.data() + (size_t)omp_get_thread_num() * NP;
        R* g1 = gw; R* g2 = g1 + 64 * 32; R* g3 = g2 + 16 * 64; R* g4 = g3 + 64 * 32; R* g5 = g4 + 64 * 64;
#pragma omp for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
            const uint32_t i = (uint32_t)ii;
            R x[32], h1[64], s[16], xc[32], h2[64], h3[64], y[3];
            for (int k = 0; k < 32; ++k) x[k] = feat[(size_t)k * B + i];
            for (int r = 0; r < 64; ++r) { R a = 0; for (int k = 0; k < 32; ++k) a += W1[r * 32 + k] * x[k]; h1[r] = a > 0 ? a : 0; }
            for (int r = 0; r < 16; ++r) { R a = 0; for (int k = 0; k < 64; ++k) a += W2[r * 64 + k] * h1[k]; s[r] = a; }
            double shd[16];
            const float d3[3] = {dir01[i], dir01[i + B], dir01[i + 2 * B]};
            oc_sh4_f64(1, 1, d3, shd);
            for (int k = 0; k < 16; ++k) { xc[k] = s[k]; xc[16 + k] = (R)shd[k]; }
            for (int r = 0; r < 64; ++r) { R a = 0; for (int k = 0; k < 32; ++k) a += W3[r * 32 + k] * xc[k]; h2[r] = a > 0 ? a : 0; }
            for (int r = 0; r < 64; ++r) { R a = 0; for (int k = 0; k < 64; ++k) a += W4[r * 64 + k] * h2[k]; h3[r] = a > 0 ? a : 0; }
            for (int r = 0; r < 3; ++r) { R a = 0; for (int k = 0; k < 64; ++k) a += W5[r * 64 + k] * h3[k]; y[r] = (R)1 / ((R)1 + std::exp(-a)); }
            if (dens_out) dens_out[i] = s[0];
            if (rgb_out) for (int r = 0; r < 3; ++r) rgb_out[(size_t)r * B + i] = y[r];
            if (!G) continue;  // forward only
            // backward
            R dO[3], dH3[64], dH2[64], dS[16], dH1[64];
            for (int r = 0; r < 3; ++r) dO[r] = G[(size_t)r * B + i] * y[r] * ((R)1 - y[r]);
            for (int r = 0; r < 3; ++r) for (int k = 0; k < 64; ++k) g5[r * 64 + k] += dO[r] * h3[k];
            for (int k = 0; k < 64; ++k) { R a = 0; for (int r = 0; r < 3; ++r) a += W5[r * 64 + k] * dO[r]; dH3[k] = h3[k] > 0 ? a : 0; }
            for (int r = 0; r < 64; ++r) for (int k = 0; k < 64; ++k) g4[r * 64 + k] += dH3[r] * h2[k];
            for (int k = 0; k < 64; ++k) { R a = 0; for (int r = 0; r < 64; ++r) a += W4[r * 64 + k] * dH3[r]; dH2[k] = h2[k] > 0 ? a : 0; }
            for (int r = 0; r < 64; ++r) for (int k = 0; k < 32; ++k) g3[r * 32 + k] += dH2[r] * xc[k];
            for (int k = 0; k < 16; ++k) { R a = 0; for (int r = 0; r < 64; ++r) a += W3[r * 32 + k] * dH2[r]; dS[k] = a; }
            R g0 = dS[0] + loss_scale * (R)dL_ds[i];
            if (extra) g0 += loss_scale * (R)extra[i];
            if (t[i] < 1.0f) g0 *= (R)t[i] * (R)t[i];
            dS[0] = g0;
            for (int r = 0; r < 16; ++r) for (int k = 0; k < 64; ++k) g2[r * 64 + k] += dS[r] * h1[k];
            for (int k = 0; k < 64; ++k) { R a = 0; for (int r = 0; r < 16; ++r) a += W2[r * 64 + k] * dS[r]; dH1[k] = h1[k] > 0 ? a : 0; }
            for (int r = 0; r < 64; ++r) for (int k = 0; k < 32; ++k) g1[r * 32 + k] += dH1[r] * x[k];
            for (int k = 0; k < 32; ++k) { R a = 0; for (int r = 0; r < 64; ++r) a += W1[r * 32 + k] * dH1[r]; dfeat[(size_t)k * B + i] = a; }
        }
    }
    if (G)
        for (int th = 0; th < nt; ++th)
            for (int e = 0; e < NP; ++e) dW[e] += part[(size_t)th * NP + e];
}

extern "C" void oc_mlp_train_f32(uint32_t n, uint32_t B, const float* net, const float* feat, const float* dir01,
    const float* G, const float* dL_ds, const float* extra, const float* t, float loss_scale, float* dens_out,
    float* rgb_out, float* dfeat, float* dW) {
    mlp_train<float>(n, B, net, feat, dir01, G, dL_ds, extra, t, loss_scale, dens_out, rgb_out, dfeat, dW);
}

extern "C" void oc_mlp_train_f64(uint32_t n, uint32_t B, const double* net, const double* feat, const float* dir01,
    const double* G, const float* dL_ds, const float* extra, const float* t, double loss_scale, double* dens_out,
    double* rgb_out, double* dfeat, double* dW) {
    mlp_train<double>(n, B, net, feat, dir01, G, dL_ds, extra, t, loss_scale, dens_out, rgb_out, dfeat, dW);
}

/* ===================================================================== */
/* Analytic Gaussian-blob field used for synthetic scenes. blob k occupies */
/* 8 floats: centre xyz (world), sigma s, peak p, rgb.                     */
/* density(x) = sum_k p_k exp(-|x-c_k|^2 / (2 s_k^2)); colour = density-   */
/* weighted mean of the blob colours.                                      */
/* ===================================================================== */
static inline void blob_eval_d(const float* xyz3, uint32_t nb, const float* bl, double& sig, double rgb[3]) {
    double s = 0, r = 0, g = 0, b = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        const float* q = bl + 8 * k;
        const double dx = (double)xyz3[0] - q[0], dy = (double)xyz3[1] - q[1], dz = (double)xyz3[2] - q[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        const double rho = (double)q[4] * std::exp(-d2 / (2.0 * (double)q[3] * (double)q[3]));
        s += rho; r += rho * q[5]; g += rho * q[6]; b += rho * q[7];
    }
    sig = s;
    const double inv = s > 0 ? 1.0 / s : 0.0;
    rgb[0] = r * inv; rgb[1] = g * inv; rgb[2] = b * inv;
}

extern "C" void oc_blob_eval(uint32_t n, const float* xyz, uint32_t nb, const float* bl, float* sigma, float* rgb) {
    for (uint32_t i = 0; i < n; ++i) {
        double s, c[3];
        blob_eval_d(xyz + 3 * i, nb, bl, s, c);
        sigma[i] = (float)s;
        rgb[3 * i] = (float)c[0]; rgb[3 * i + 1] = (float)c[1]; rgb[3 * i + 2] = (float)c[2];
    }
}

/* Front-to-back compositing of the analytic field over marched samples,
 * with the reference's early termination (training-kernels-fused.cuh:243-258)
 * and its render rule that a sample's colour counts only when its alpha
 * exceeds 1e-4 (rendering-kernels.cu:570-578). Output rgba[4][n_rays]:
 * premultiplied rgb and accumulated alpha. */
extern "C" void oc_composite_analytic_f64(uint32_t n_rays, const uint32_t* steps, const uint32_t* offs,
    const float* pos, uint32_t B, float aabb, const float* sdt, uint32_t nb, const float* bl, double min_T,
    double* rgba) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        double r = 0, g = 0, b = 0, a = 0, T = 1;
        for (uint32_t k = 0; k < steps[i]; ++k) {
            const uint32_t s = offs[i] + k;
            const float xw[3] = {(pos[s] - 0.5f) * aabb, (pos[s + B] - 0.5f) * aabb, (pos[s + 2 * B] - 0.5f) * aabb};
            double sig, c[3];
            blob_eval_d(xw, nb, bl, sig, c);
            const double alpha = 1.0 - std::exp(-sig * (double)sdt[s]);
            const double w = T * alpha;
            // colour only when alpha > 1e-4 (rendering-kernels.cu:570-578)
            if (alpha > 1e-4) { r += w * c[0]; g += w * c[1]; b += w * c[2]; }
            a += w;
            T *= 1.0 - alpha;
            if (T < min_T) break;
        }
        rgba[i] = r; rgba[i + n_rays] = g; rgba[i + 2 * n_rays] = b; rgba[i + 3 * n_rays] = a;
    }
}

/* Render rays of pixels [start, start + n) for one NeRF with an identity
 * transform: generate_linear_buffer_of_rays_kernel + fill_ray_buffers
 * (linear-buffer-ray-batch-coordinator.cu:11-34, ray-batch-coordinator.cuh:46-84)
 * then prepare_for_linear_raymarching_kernel (rendering-kernels.cu:174-260):
 * the origin moves to the box entry max(tmin, 1e-4), t starts at 0 and
 * t_max = min(box exit, far - near) measured from the original origin. */
extern "C" void oc_render_rays(uint32_t n, uint32_t start, uint32_t B, const tn_camera_t* cam,
    const tn_bbox_t* bbox, float* ori, float* dir, float* ray_t, float* ray_tmax, uint8_t* alive) {
    const int w = cam->resolution[0];
    for (uint32_t i = 0; i < n; ++i) {
        const int ii = (int)(i + start);
        const int iy = ii / w, ix = ii - iy * w;
        f3 o, d;
        camera_ray(*cam, (float)ix, (float)iy, o, d);
        float tmax = cam->far_plane - cam->near_plane;
        const f3 id = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
        const float vol = (bbox->max_x - bbox->min_x) * (bbox->max_y - bbox->min_y) * (bbox->max_z - bbox->min_z);
        float t0, t1;
        ray_t[i] = 0.0f;
        if (!(vol > 0.0f) || !bbox_intersect(*bbox, o, id, t0, t1)) {
            alive[i] = 0;
            ori[i] = o.x; ori[i + B] = o.y; ori[i + 2 * B] = o.z;
            dir[i] = d.x; dir[i + B] = d.y; dir[i + 2 * B] = d.z;
            ray_tmax[i] = tmax;
            continue;
        }
        t0 = fmaxf(t0, 1e-4f);
        alive[i] = 1;
        ray_tmax[i] = fminf(t1, tmax);
        ori[i] = o.x + t0 * d.x; ori[i + B] = o.y + t0 * d.y; ori[i + 2 * B] = o.z + t0 * d.z;
        dir[i] = d.x; dir[i + B] = d.y; dir[i + 2 * B] = d.z;
    }
}

/* The reference's render loop for one NeRF, restated over the analytic
 * field (rendering-kernels.cu:264-330 march, :404-478 occupancy filter,
 * :480-528 accumulate, :530-618 composite; renderer.cu:215,305 steps). The
 * ray starts at its box entry (t = 0) and steps by get_dt(t, cone, dt_min,
 * dt_max) with no empty-space skipping; a sample counts when it is inside
 * the box and its cell (cascade chosen by dt) is occupied; alpha =
 * 1 - exp(-sigma dt), colour counted only when alpha > 1e-4 (the reference
 * divides the accumulated colour by alpha with that guard); stop once T <
 * min_T, past t_max or after max_steps steps. fp64 accumulation. */
extern "C" void oc_render_fixed_step_analytic_f64(uint32_t n_rays, uint32_t B, const float* ori, const float* dir,
    const float* ray_t, const float* ray_tmax, const tn_bbox_t* bbox, float inv_aabb, const uint8_t* bits,
    int n_levels, int res, float cone, float dt_min, float dt_max, uint32_t max_steps, uint32_t nb, const float* bl,
    double min_T, double* rgba) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t ii = 0; ii < (int64_t)n_rays; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float ox = ori[i], oy = ori[i + B], oz = ori[i + 2 * B];
        const float dx = dir[i], dy = dir[i + B], dz = dir[i + 2 * B];
        double r = 0, g = 0, b = 0, a = 0, T = 1;
        float t = ray_t[i];
        for (uint32_t k = 0; k < max_steps && t <= ray_tmax[i]; ++k) {
            const float x = ox + t * dx, y = oy + t * dy, z = oz + t * dz;
            const float dt = get_dt(t, cone, dt_min, dt_max);
            if (bbox_contains(*bbox, x, y, z)) {
                // the network step is normalised by the scene size
                // (rendering-kernels.cu:445); it picks the cascade (:447) and
                // turns density into alpha (:517-520)
                const float ndt = inv_aabb * dt;
                const int lvl = oc_grid_level_at(n_levels, res, x, y, z, ndt);
                if (grid_occupied(bits, res, lvl, x, y, z)) {
                    const float xw[3] = {x, y, z};
                    double sig, c[3];
                    blob_eval_d(xw, nb, bl, sig, c);
                    const double alpha = 1.0 - std::exp(-sig * (double)ndt);
                    const double w = T * alpha;
                    if (alpha > 1e-4) { r += w * c[0]; g += w * c[1]; b += w * c[2]; }
                    a += w;
                    T *= 1.0 - alpha;
                    if (T < min_T) break;
                }
            }
            t += dt;
        }
        rgba[i] = r; rgba[i + n_rays] = g; rgba[i + 2 * n_rays] = b; rgba[i + 3 * n_rays] = a;
    }
}

extern "C" int oc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
