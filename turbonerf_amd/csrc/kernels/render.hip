// Render-side ray kernels: row-major pixel rays (reference
// generate_linear_buffer_of_rays_kernel + fill_ray_buffers,
// src/rendering/ray-batch-coordinators/linear-buffer-ray-batch-coordinator.cu:11-34,
// ray-batch-coordinator.cuh:46-84), ray preparation against a NeRF's render
// box in its local frame (reference prepare_for_linear_raymarching_kernel,
// src/utils/rendering-kernels.cu:175-260) and the SoA -> interleaved RGBA
// repack into the output image (reference copy_packed_rgba_linear_buffer_kernel,
// linear-buffer-ray-batch-coordinator.cu:54-79).
#pragma clang fp contract(off)

#include "camera.hpp"
#include "common.hpp"

namespace tn {

// Pixel rays for pixels [start, start + n) of the camera image, transformed
// by `to_local` (3x4, world -> NeRF frame), clipped to `bbox` there: the
// origin moves to the entry point (t = 0 there) and t_max is the exit
// distance (capped by far - near). Rays that miss are not alive.
__global__ void render_rays_kernel(uint32_t n, uint32_t start, uint32_t B, tn_camera_t cam, tn_bbox_t bbox,
                                   const float* __restrict__ to_local_ptr, float* __restrict__ ori,
                                   float* __restrict__ dir, float* __restrict__ ray_t, float* __restrict__ ray_tmax,
                                   uint32_t* __restrict__ pixel, uint8_t* __restrict__ alive) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = start + i;
    const uint32_t w = (uint32_t)cam.resolution[0];
    const uint32_t iy = p / w, ix = p - iy * w;
    float ox, oy, oz, dx, dy, dz;
    camera_ray(cam, (float)ix, (float)iy, ox, oy, oz, dx, dy, dz);
    float M[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) M[k] = to_local_ptr ? to_local_ptr[k] : ((k == 0 || k == 5 || k == 10) ? 1.0f : 0.0f);
    const float lox = M[0] * ox + M[1] * oy + M[2] * oz + M[3];
    const float loy = M[4] * ox + M[5] * oy + M[6] * oz + M[7];
    const float loz = M[8] * ox + M[9] * oy + M[10] * oz + M[11];
    float ldx = M[0] * dx + M[1] * dy + M[2] * dz;
    float ldy = M[4] * dx + M[5] * dy + M[6] * dz;
    float ldz = M[8] * dx + M[9] * dy + M[10] * dz;
    const float ll = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
    ldx = ldx / ll;
    ldy = ldy / ll;
    ldz = ldz / ll;
    pixel[i] = p;
    float tmin, tmax;
    const bool vol = (bbox.max_x - bbox.min_x) * (bbox.max_y - bbox.min_y) * (bbox.max_z - bbox.min_z) > 0.0f;
    if (!vol || !bbox_intersect(bbox, lox, loy, loz, 1.0f / ldx, 1.0f / ldy, 1.0f / ldz, tmin, tmax) || tmax < 0.0f) {
        alive[i] = 0;
        return;
    }
    tmin = fmaxf(tmin, 1e-4f);
    const float far_t = cam.far_plane - cam.near_plane;
    // the entry point, clamped into the box: rounding can leave it an ulp
    // outside, where the march would end the ray before its first sample
    ori[i] = fminf(fmaxf(lox + tmin * ldx, bbox.min_x), bbox.max_x);
    ori[i + B] = fminf(fmaxf(loy + tmin * ldy, bbox.min_y), bbox.max_y);
    ori[i + 2 * B] = fminf(fmaxf(loz + tmin * ldz, bbox.min_z), bbox.max_z);
    dir[i] = ldx;
    dir[i + B] = ldy;
    dir[i + 2 * B] = ldz;
    ray_t[i] = 0.0f;
    ray_tmax[i] = fminf(tmax, far_t) - tmin;
    alive[i] = ray_tmax[i] > 0.0f ? 1 : 0;
}

// out[4*(pixel)+c] = rgba[c][i] for the rays [0, n); rays that were not kept
// in the batch are left untouched
__global__ void write_rgba_kernel(uint32_t n, const uint32_t* __restrict__ info, uint32_t RB,
                                  const uint32_t* __restrict__ pixel, const uint8_t* __restrict__ alive,
                                  const float* __restrict__ rgba, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = info ? min(info[0], n) : n;
    if (i >= m) return;
    const uint32_t p = pixel[i];
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (alive[i]) v = make_float4(rgba[i], rgba[i + RB], rgba[i + 2 * RB], rgba[i + 3 * RB]);
    reinterpret_cast<float4*>(out)[p] = v;
}

// ---------------------------------------------------------------------------
// Several NeRFs in one render (reference renderer.cu:344-521,
// rendering-kernels.cu:175-260,264-332,336-528,531-606): every ray marches in
// world space with the fixed step rule, all NeRFs evaluate the same samples,
// and a sample's alphas add (capped at 1) while its colour is their
// alpha-weighted mean. Samples are laid out step-major, sample s of ray i at
// s * n_rays + i, k steps per ray per pass.
// ---------------------------------------------------------------------------

TN_DEVICE void xform(const float* M, float x, float y, float z, float& ox, float& oy, float& oz) {
    ox = M[0] * x + M[1] * y + M[2] * z + M[3];
    oy = M[4] * x + M[5] * y + M[6] * z + M[7];
    oz = M[8] * x + M[9] * y + M[10] * z + M[11];
}
TN_DEVICE void xform_dir(const float* M, float x, float y, float z, float& ox, float& oy, float& oz) {
    ox = M[0] * x + M[1] * y + M[2] * z;
    oy = M[4] * x + M[5] * y + M[6] * z;
    oz = M[8] * x + M[9] * y + M[10] * z;
}

// world rays of pixels [start, start + n): the origin moves to the first
// entry into any NeRF's box (t = 0 there), t_max is the last exit (capped by
// far - near), rays that miss every box are not alive
__global__ void render_rays_multi_kernel(uint32_t n, uint32_t start, uint32_t B, tn_camera_t cam, uint32_t n_nerfs,
                                         const tn_render_nerf_t* __restrict__ nerfs, float* __restrict__ ori,
                                         float* __restrict__ dir, float* __restrict__ ray_t,
                                         float* __restrict__ ray_tmax, uint32_t* __restrict__ steps,
                                         uint32_t* __restrict__ pixel, uint8_t* __restrict__ alive,
                                         float* __restrict__ trans) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = start + i;
    const uint32_t w = (uint32_t)cam.resolution[0];
    const uint32_t iy = p / w, ix = p - iy * w;
    float ox, oy, oz, dx, dy, dz;
    camera_ray(cam, (float)ix, (float)iy, ox, oy, oz, dx, dy, dz);
    pixel[i] = p;
    steps[i] = 0;
    trans[i] = 1.0f;
    float t_min = 3.402823466e38f, t_max = 0.0f;
    bool hit = false;
    for (uint32_t k = 0; k < n_nerfs; ++k) {
        const tn_render_nerf_t& nf = nerfs[k];
        const tn_bbox_t& b = nf.render_box;
        if (!((b.max_x - b.min_x) * (b.max_y - b.min_y) * (b.max_z - b.min_z) > 0.0f)) continue;
        float lox, loy, loz, ldx, ldy, ldz;
        xform(nf.to_local, ox, oy, oz, lox, loy, loz);
        xform_dir(nf.to_local, dx, dy, dz, ldx, ldy, ldz);
        float t0, t1;
        if (!bbox_intersect(b, lox, loy, loz, 1.0f / ldx, 1.0f / ldy, 1.0f / ldz, t0, t1) || t1 < 0.0f) continue;
        hit = true;
        t0 = fmaxf(t0, 1e-4f);
        t_min = fminf(t_min, t0);
        t_max = fmaxf(t_max, t1);
    }
    if (!hit) {
        alive[i] = 0;
        return;
    }
    ori[i] = ox + t_min * dx;
    ori[i + B] = oy + t_min * dy;
    ori[i + 2 * B] = oz + t_min * dz;
    dir[i] = dx;
    dir[i + B] = dy;
    dir[i + 2 * B] = dz;
    ray_t[i] = 0.0f;
    ray_tmax[i] = fminf(t_max, cam.far_plane - cam.near_plane) - t_min;
    alive[i] = ray_tmax[i] > 0.0f ? 1 : 0;
}

// up to k fixed steps of each live ray from its t (dt = clamp(t cone, dt_min,
// dt_max)); a step past t_max or past max_steps is absent (dt < 0)
__global__ void render_march_fixed_kernel(uint32_t n, uint32_t k, uint32_t S, float cone, float dt_min, float dt_max,
                                          uint32_t max_steps, const uint8_t* __restrict__ alive,
                                          const float* __restrict__ ori, const float* __restrict__ dir, uint32_t B,
                                          float* __restrict__ ray_t, const float* __restrict__ ray_tmax,
                                          uint32_t* __restrict__ steps, float* __restrict__ s_pos,
                                          float* __restrict__ s_dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool live = alive[i] != 0;
    float t = ray_t[i];
    const float tmax = ray_tmax[i];
    uint32_t st = steps[i];
    const float ox = ori[i], oy = ori[i + B], oz = ori[i + 2 * B];
    const float dx = dir[i], dy = dir[i + B], dz = dir[i + 2 * B];
    for (uint32_t s = 0; s < k; ++s) {
        const uint32_t q = s * n + i;
        if (!live || t > tmax || st >= max_steps) {
            s_dt[q] = -1.0f;
            continue;
        }
        s_pos[q] = ox + t * dx;
        s_pos[q + S] = oy + t * dy;
        s_pos[q + 2 * S] = oz + t * dz;
        const float dt = get_dt(t, cone, dt_min, dt_max);
        s_dt[q] = dt;
        t += dt;
        ++st;
    }
    ray_t[i] = t;
    steps[i] = st;
}

// the samples one NeRF evaluates: inside its box (in its frame), folded by
// its repeater, on an occupied cell of the cascade the normalised step picks;
// valid[q] = 1 and the unit-box network inputs (position, direction (0.5 d +
// 0.5 of the unit local direction), normalised step) for those
__global__ void render_localize_kernel(uint32_t n, uint32_t k, uint32_t S, tn_render_nerf_t nf,
                                       tn_occupancy_grid_t grid, const float* __restrict__ dir, uint32_t B,
                                       const float* __restrict__ s_pos, const float* __restrict__ s_dt,
                                       uint32_t* __restrict__ valid, float* __restrict__ net_pos,
                                       float* __restrict__ net_dir, float* __restrict__ net_dt) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n * k) return;
    valid[q] = 0u;
    const float dtw = s_dt[q];
    if (!(dtw >= 0.0f)) return;
    const uint32_t i = q % n;
    float x, y, z;
    xform(nf.to_local, s_pos[q], s_pos[q + S], s_pos[q + 2 * S], x, y, z);
    if (!bbox_contains(nf.render_box, x, y, z)) return;
    float ldx, ldy, ldz;
    xform_dir(nf.to_local, dir[i], dir[i + B], dir[i + 2 * B], ldx, ldy, ldz);
    const float scale = sqrtf(ldx * ldx + ldy * ldy + ldz * ldz);
    if (nf.repeat) {
        const tn_bbox_t& r = nf.repeat_source;
        const float ex = r.max_x - r.min_x, ey = r.max_y - r.min_y, ez = r.max_z - r.min_z;
        x = fmodf(fmodf(x - r.min_x, ex) + ex, ex) + r.min_x;
        y = fmodf(fmodf(y - r.min_y, ey) + ey, ey) + r.min_y;
        z = fmodf(fmodf(z - r.min_z, ez) + ez, ez) + r.min_z;
    }
    const float ndt = nf.inv_aabb * (scale * dtw);
    const float inv_res = 1.0f / (float)grid.resolution;
    const int level = grid_level_at(grid.n_levels, inv_res, x, y, z, ndt);
    const uint32_t m = voxel_morton_index((float)grid.resolution, (uint32_t)grid.resolution - 1u, level, x, y, z);
    if (((grid.bitfield[m] >> level) & 1u) == 0u) return;
    valid[q] = 1u;
    const float is = 1.0f / nf.training_size;
    net_pos[q] = fminf(fmaxf((x - nf.training_min[0]) * is, 0.0f), 1.0f);
    net_pos[q + S] = fminf(fmaxf((y - nf.training_min[1]) * is, 0.0f), 1.0f);
    net_pos[q + 2 * S] = fminf(fmaxf((z - nf.training_min[2]) * is, 0.0f), 1.0f);
    net_dir[q] = 0.5f * (ldx / scale) + 0.5f;
    net_dir[q + S] = 0.5f * (ldy / scale) + 0.5f;
    net_dir[q + 2 * S] = 0.5f * (ldz / scale) + 0.5f;
    net_dt[q] = ndt;
}

// the valid samples' inputs packed to the front, in sample order (offsets:
// exclusive scan of valid)
__global__ void render_compact_kernel(uint32_t nq, uint32_t S, const uint32_t* __restrict__ valid,
                                      const uint32_t* __restrict__ offs, const float* __restrict__ net_pos,
                                      const float* __restrict__ net_dir, float* __restrict__ c_pos,
                                      float* __restrict__ c_dir) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq || !valid[q]) return;
    const uint32_t j = offs[q];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        c_pos[j + c * S] = net_pos[q + c * S];
        c_dir[j + c * S] = net_dir[q + c * S];
    }
}

// a NeRF's outputs added into the samples: alpha = 1 - exp(-sigma dt)
// (sigma = exp(raw) for a network, the field's sigma for a synthetic one),
// rgba += (c alpha, alpha), count += 1
template <typename T>
__global__ void render_accumulate_kernel(uint32_t nq, uint32_t S, const uint32_t* __restrict__ valid,
                                         const uint32_t* __restrict__ offs, const T* __restrict__ density,
                                         const T* __restrict__ rgb, int raw_density, const float* __restrict__ net_dt,
                                         float* __restrict__ s_rgba, uint32_t* __restrict__ count) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq || !valid[q]) return;
    const uint32_t j = offs[q];
    const float d = (float)density[j];
    const float sigma = raw_density ? __expf(d) : d;
    const float a = 1.0f - __expf(-sigma * net_dt[q]);
    s_rgba[q] += (float)rgb[j] * a;
    s_rgba[q + S] += (float)rgb[j + S] * a;
    s_rgba[q + 2 * S] += (float)rgb[j + 2 * S] * a;
    s_rgba[q + 3 * S] += a;
    count[q] += 1u;
}

// front-to-back compositing of each ray's k samples of this pass into its
// pixel (interleaved RGBA); the ray ends when its transmittance falls below
// min_T, at its last step, or at an absent step
__global__ void render_composite_fixed_kernel(uint32_t n, uint32_t k, uint32_t S, float min_T,
                                              const float* __restrict__ s_dt, const float* __restrict__ s_rgba,
                                              const uint32_t* __restrict__ count, const uint32_t* __restrict__ pixel,
                                              uint8_t* __restrict__ alive, float* __restrict__ trans,
                                              float* __restrict__ image) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !alive[i]) return;
    float T = trans[i], r = 0.0f, g = 0.0f, b = 0.0f, acc = 0.0f;
    bool done = false;
    for (uint32_t s = 0; s < k; ++s) {
        const uint32_t q = s * n + i;
        if (!(s_dt[q] >= 0.0f)) {
            done = true;
            break;
        }
        if (count[q] == 0u) continue;
        const float sa = s_rgba[q + 3 * S];
        const float inv = sa > 1e-4f ? 1.0f / sa : 0.0f;
        const float a = fminf(sa, 1.0f);
        const float wgt = a * T;
        r += wgt * (s_rgba[q] * inv);
        g += wgt * (s_rgba[q + S] * inv);
        b += wgt * (s_rgba[q + 2 * S] * inv);
        acc += wgt;
        T *= 1.0f - a;
        if (T < min_T) {
            done = true;
            break;
        }
    }
    trans[i] = T;
    float4* px = reinterpret_cast<float4*>(image) + pixel[i];
    float4 v = *px;
    v.x += r;
    v.y += g;
    v.z += b;
    v.w += acc;
    *px = v;
    if (done) alive[i] = 0;
}

__global__ void count_nonzero_kernel(uint32_t n, const uint8_t* __restrict__ v, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t b = __ballot(i < n && v[i] != 0);
    if (__lane_id() == 0 && b) atomicAdd(out, (uint32_t)__popcll(b));
}

__global__ void fill_kernel(uint32_t n, uint32_t* __restrict__ p, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_render_rays_multi(uint32_t n_rays, uint32_t start_pixel, uint32_t stride, tn_camera_t camera,
                                    uint32_t n_nerfs, const tn_render_nerf_t* nerfs, float* ori, float* dir,
                                    float* ray_t, float* ray_tmax, uint32_t* ray_steps, uint32_t* pixel,
                                    uint8_t* alive, float* trans, hipStream_t stream) {
    TN_CHECK_ARG(nerfs && ori && dir && ray_t && ray_tmax && ray_steps && pixel && alive && trans);
    TN_CHECK_ARG(n_rays <= stride && camera.resolution[0] > 0 && n_nerfs > 0);
    if (n_rays == 0) return TN_OK;
    render_rays_multi_kernel<<<blocks_for(n_rays, 256), 256, 0, stream>>>(
        n_rays, start_pixel, stride, camera, n_nerfs, nerfs, ori, dir, ray_t, ray_tmax, ray_steps, pixel, alive, trans);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_render_march_fixed(uint32_t n_rays, uint32_t k, uint32_t sample_stride, float cone_angle,
                                     float dt_min, float dt_max, uint32_t max_steps, const uint8_t* alive,
                                     const float* ori, const float* dir, uint32_t ray_stride, float* ray_t,
                                     const float* ray_tmax, uint32_t* ray_steps, float* s_pos, float* s_dt,
                                     hipStream_t stream) {
    TN_CHECK_ARG(alive && ori && dir && ray_t && ray_tmax && ray_steps && s_pos && s_dt && n_rays <= ray_stride);
    TN_CHECK_ARG((uint64_t)n_rays * k <= sample_stride && 3ull * sample_stride < (1ull << 32));
    if (n_rays == 0 || k == 0) return TN_OK;
    render_march_fixed_kernel<<<blocks_for(n_rays, 256), 256, 0, stream>>>(n_rays, k, sample_stride, cone_angle,
                                                                          dt_min, dt_max, max_steps, alive, ori, dir,
                                                                          ray_stride, ray_t, ray_tmax, ray_steps,
                                                                          s_pos, s_dt);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_render_localize(uint32_t n_rays, uint32_t k, uint32_t sample_stride, tn_render_nerf_t nerf,
                                  tn_occupancy_grid_t grid, const float* dir, uint32_t ray_stride, const float* s_pos,
                                  const float* s_dt, uint32_t* valid, float* net_pos, float* net_dir, float* net_dt,
                                  hipStream_t stream) {
    TN_CHECK_ARG(dir && s_pos && s_dt && valid && net_pos && net_dir && net_dt && grid.bitfield);
    TN_CHECK_ARG((uint64_t)n_rays * k <= sample_stride && n_rays <= ray_stride && nerf.training_size > 0.0f);
    if (n_rays == 0 || k == 0) return TN_OK;
    render_localize_kernel<<<blocks_for((uint64_t)n_rays * k, 256), 256, 0, stream>>>(
        n_rays, k, sample_stride, nerf, grid, dir, ray_stride, s_pos, s_dt, valid, net_pos, net_dir, net_dt);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_render_compact(uint32_t n_samples, uint32_t sample_stride, const uint32_t* valid,
                                 const uint32_t* offsets, const float* net_pos, const float* net_dir, float* c_pos,
                                 float* c_dir, hipStream_t stream) {
    TN_CHECK_ARG(valid && offsets && net_pos && net_dir && c_pos && c_dir && n_samples <= sample_stride);
    if (n_samples == 0) return TN_OK;
    render_compact_kernel<<<blocks_for(n_samples, 256), 256, 0, stream>>>(n_samples, sample_stride, valid, offsets,
                                                                         net_pos, net_dir, c_pos, c_dir);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// dtype: 0 = fp32 outputs, 1 = fp16; raw_density: 1 = density is exp(raw)
extern "C" int tn_render_accumulate(int dtype, uint32_t n_samples, uint32_t sample_stride, const uint32_t* valid,
                                    const uint32_t* offsets, const void* density, const void* rgb, int raw_density,
                                    const float* net_dt, float* s_rgba, uint32_t* count, hipStream_t stream) {
    TN_CHECK_ARG(valid && offsets && density && rgb && net_dt && s_rgba && count && n_samples <= sample_stride);
    if (n_samples == 0) return TN_OK;
    const unsigned blocks = blocks_for(n_samples, 256);
    if (dtype == 0)
        render_accumulate_kernel<float><<<blocks, 256, 0, stream>>>(n_samples, sample_stride, valid, offsets,
                                                                   static_cast<const float*>(density),
                                                                   static_cast<const float*>(rgb), raw_density,
                                                                   net_dt, s_rgba, count);
    else if (dtype == 1)
        render_accumulate_kernel<_Float16><<<blocks, 256, 0, stream>>>(n_samples, sample_stride, valid, offsets,
                                                                      static_cast<const _Float16*>(density),
                                                                      static_cast<const _Float16*>(rgb), raw_density,
                                                                      net_dt, s_rgba, count);
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_render_composite_fixed(uint32_t n_rays, uint32_t k, uint32_t sample_stride, float min_transmittance,
                                         const float* s_dt, const float* s_rgba, const uint32_t* count,
                                         const uint32_t* pixel, uint8_t* alive, float* trans, float* image,
                                         hipStream_t stream) {
    TN_CHECK_ARG(s_dt && s_rgba && count && pixel && alive && trans && image);
    TN_CHECK_ARG((uint64_t)n_rays * k <= sample_stride);
    if (n_rays == 0 || k == 0) return TN_OK;
    render_composite_fixed_kernel<<<blocks_for(n_rays, 256), 256, 0, stream>>>(
        n_rays, k, sample_stride, min_transmittance, s_dt, s_rgba, count, pixel, alive, trans, image);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// Preview expansion (reference copy_packed_rgba_rectangular_grid_kernel,
// rectangular-grid-ray-batch-coordinator.cu:62-105): output pixel (x, y)
// takes grid pixel (x gw / W, y gh / H).
__global__ void expand_nearest_kernel(uint32_t gw, uint32_t gh, uint32_t W, uint32_t H, const float* __restrict__ in,
                                      float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W * H) return;
    const uint32_t y = i / W, x = i - y * W;
    const uint32_t gx = (uint32_t)(((uint64_t)x * gw) / W), gy = (uint32_t)(((uint64_t)y * gh) / H);
    reinterpret_cast<float4*>(out)[i] = reinterpret_cast<const float4*>(in)[gy * gw + gx];
}

extern "C" int tn_render_expand_nearest(uint32_t grid_w, uint32_t grid_h, uint32_t width, uint32_t height,
                                        const float* grid_rgba, float* rgba, hipStream_t stream) {
    TN_CHECK_ARG(grid_rgba && rgba && grid_w > 0 && grid_h > 0 && (uint64_t)width * height < (1ull << 32));
    if (width == 0 || height == 0) return TN_OK;
    expand_nearest_kernel<<<blocks_for((uint64_t)width * height, 256), 256, 0, stream>>>(grid_w, grid_h, width, height,
                                                                                       grid_rgba, rgba);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// *out = number of non-zero bytes of v[0, n) (out is cleared first)
extern "C" int tn_count_nonzero(uint32_t n, const uint8_t* v, uint32_t* out, hipStream_t stream) {
    TN_CHECK_ARG(v && out);
    fill_kernel<<<1, 64, 0, stream>>>(1, out, 0u);
    if (n) count_nonzero_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(n, v, out);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// p[0, n) = v
extern "C" int tn_fill_u32(uint32_t n, uint32_t* p, uint32_t v, hipStream_t stream) {
    TN_CHECK_ARG(p);
    if (n == 0) return TN_OK;
    fill_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(n, p, v);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// to_local: optional device pointer to 12 floats (world -> NeRF frame).
extern "C" int tn_render_rays(uint32_t n_rays, uint32_t start_pixel, uint32_t stride, tn_camera_t camera,
                              tn_bbox_t render_bbox, const float* to_local, float* ori, float* dir, float* ray_t,
                              float* ray_tmax, uint32_t* pixel, uint8_t* alive, hipStream_t stream) {
    TN_CHECK_ARG(ori && dir && ray_t && ray_tmax && pixel && alive && n_rays <= stride && camera.resolution[0] > 0);
    if (n_rays == 0) return TN_OK;
    render_rays_kernel<<<blocks_for(n_rays, 256), 256, 0, stream>>>(n_rays, start_pixel, stride, camera, render_bbox,
                                                                    to_local, ori, dir, ray_t, ray_tmax, pixel, alive);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// rgba: [4][rgba_stride] premultiplied per ray; out: interleaved RGBA float image.
extern "C" int tn_render_write_rgba(uint32_t n_rays, const uint32_t* info, uint32_t rgba_stride, const uint32_t* pixel,
                                    const uint8_t* alive, const float* rgba, float* out, hipStream_t stream) {
    TN_CHECK_ARG(pixel && alive && rgba && out && n_rays <= rgba_stride);
    if (n_rays == 0) return TN_OK;
    write_rgba_kernel<<<blocks_for(n_rays, 256), 256, 0, stream>>>(n_rays, info, rgba_stride, pixel, alive, rgba, out);
    TN_CHECK_LAUNCH();
    return TN_OK;
}
