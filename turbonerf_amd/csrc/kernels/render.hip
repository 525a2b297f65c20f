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

}  // namespace tn

using namespace tn;

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
