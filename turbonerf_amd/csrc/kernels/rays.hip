// Training-ray generation: seeded uniforms and camera rays clipped to the
// training box (reference src/utils/training-batch-kernels.cuh:82-187,
// src/models/camera.cuh:94-141, src/core/trainer.cu:45-62).
#pragma clang fp contract(off)

#include "camera.hpp"
#include "common.hpp"

namespace tn {

// out[j] = uniform element (offset + j) of the stream keyed by seed.
// One thread produces four consecutive elements from one Philox block.
__global__ void uniform_kernel(uint64_t n, uint32_t k0, uint32_t k1, uint64_t offset, float* __restrict__ out) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t first_block = offset >> 2;
    const uint64_t block = first_block + tid;
    uint32_t c[4] = {(uint32_t)block, (uint32_t)(block >> 32), 0u, 0u};
    philox4x32_10(c, k0, k1);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint64_t g = block * 4 + w;
        if (g < offset) continue;
        const uint64_t j = g - offset;
        if (j < n) out[j] = u32_to_unit(c[w]);
    }
}

// The same stream at an offset read on the device (a captured step replays
// with the offset advancing on the device); the grid covers any alignment.
__global__ void uniform_dev_kernel(uint64_t n, uint32_t k0, uint32_t k1, const unsigned long long* __restrict__ d_off,
                                   float* __restrict__ out) {
    const uint64_t offset = *d_off;
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t block = (offset >> 2) + tid;
    if (block > (offset + n - 1) >> 2) return;
    uint32_t c[4] = {(uint32_t)block, (uint32_t)(block >> 32), 0u, 0u};
    philox4x32_10(c, k0, k1);
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        const uint64_t g = block * 4 + w;
        if (g < offset) continue;
        const uint64_t j = g - offset;
        if (j < n) out[j] = u32_to_unit(c[w]);
    }
}

__global__ void counter_add_kernel(unsigned long long* __restrict__ c, unsigned long long inc) { *c += inc; }

__global__ void init_training_rays_kernel(uint32_t n_rays_max, const uint32_t* __restrict__ d_n_rays,
                                          uint32_t B, uint32_t n_images, uint32_t n_pix, int32_t img_w,
                                          float n_rays_per_image, const tn_bbox_t* __restrict__ bbox,
                                          const tn_camera_t* __restrict__ cams,
                                          const uint8_t* __restrict__ images, const float* __restrict__ random,
                                          float* __restrict__ pix_rgba, uint32_t* __restrict__ ray_img_id,
                                          float* __restrict__ ori, float* __restrict__ dir,
                                          float* __restrict__ ray_t, float* __restrict__ ray_tmax,
                                          uint8_t* __restrict__ ray_alive) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_rays = d_n_rays ? min(*d_n_rays, n_rays_max) : n_rays_max;
    if (i >= n_rays) return;
    // rays are split evenly over the images (reference trainer.cu:60); a
    // non-positive n_rays_per_image asks for that split of the device count
    const float nrpi = n_rays_per_image > 0.0f ? n_rays_per_image : (float)n_rays / (float)n_images;

    uint32_t image_idx = (uint32_t)((float)i / nrpi);
    if (image_idx >= n_images) image_idx = n_images - 1;
    const float p = (float)n_pix * random[i];
    const float w = (float)img_w;
    const float x = fmodf(p, w);
    const float y = p / w;  // not floored, as in the reference (its quirk Q5)

    const tn_camera_t cam = cams[image_idx];
    float ox, oy, oz, dx, dy, dz;
    camera_ray(cam, x, y, ox, oy, oz, dx, dy, dz);
    const float idx = 1.0f / dx, idy = 1.0f / dy, idz = 1.0f / dz;
    float tmin, tmax;
    const tn_bbox_t bb = *bbox;
    // a ray that misses the box has no samples; its target is cleared so the
    // reported loss does not pick up a pixel left over from an earlier batch
    // (the reference leaves the stale value in place; gradients are the same)
    bool hit = bbox_intersect(bb, ox, oy, oz, idx, idy, idz, tmin, tmax);
    const float t_max = fminf(cam.far_plane - cam.near_plane, tmax);
    const float t = fmaxf(0.0f, tmin + 1e-5f);
    hit = hit && !(t_max < t);
    if (!hit) {
        ray_alive[i] = 0;
        pix_rgba[i] = 0.0f;
        pix_rgba[i + B] = 0.0f;
        pix_rgba[i + 2 * B] = 0.0f;
        pix_rgba[i + 3 * B] = 0.0f;
        return;
    }
    const uint64_t lp = min((uint64_t)p, (uint64_t)n_pix - 1);
    const uchar4 px = reinterpret_cast<const uchar4*>(images)[(uint64_t)n_pix * image_idx + lp];
    pix_rgba[i] = (float)px.x / 255.0f;
    pix_rgba[i + B] = (float)px.y / 255.0f;
    pix_rgba[i + 2 * B] = (float)px.z / 255.0f;
    pix_rgba[i + 3 * B] = (float)px.w / 255.0f;
    ori[i] = ox;
    ori[i + B] = oy;
    ori[i + 2 * B] = oz;
    dir[i] = dx;
    dir[i + B] = dy;
    dir[i + 2 * B] = dz;
    ray_t[i] = t;
    ray_tmax[i] = t_max;
    ray_alive[i] = 1;
    ray_img_id[i] = image_idx;
}

// Rays below the alpha threshold are kept only with the given probability
// (reference training-batch-kernels.cuh:189-208).
__global__ void deactivate_rays_kernel(uint32_t n_rays_max, const uint32_t* __restrict__ d_n_rays, uint32_t B,
                                       float alpha_threshold, float selection_probability,
                                       const float* __restrict__ pix_rgba, const float* __restrict__ random,
                                       uint8_t* __restrict__ ray_alive) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_rays = d_n_rays ? min(*d_n_rays, n_rays_max) : n_rays_max;
    if (i >= n_rays) return;
    if (pix_rgba[3 * B + i] < alpha_threshold && random[i] > selection_probability) ray_alive[i] = 0;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_fill_uniform(uint64_t n, uint64_t seed, uint64_t offset, float* out, hipStream_t stream) {
    TN_CHECK_ARG(out != nullptr);
    if (n == 0) return TN_OK;
    const uint64_t first_block = offset >> 2;
    const uint64_t last_block = (offset + n - 1) >> 2;
    const uint64_t n_threads = last_block - first_block + 1;
    uniform_kernel<<<blocks_for(n_threads, 256), 256, 0, stream>>>(n, (uint32_t)seed, (uint32_t)(seed >> 32),
                                                                   offset, out);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_fill_uniform_dev(uint64_t n, uint64_t seed, unsigned long long* d_offset, float* out,
                                   hipStream_t stream) {
    TN_CHECK_ARG(out != nullptr && d_offset != nullptr);
    if (n == 0) return TN_OK;
    uniform_dev_kernel<<<blocks_for(n / 4 + 2, 256), 256, 0, stream>>>(n, (uint32_t)seed, (uint32_t)(seed >> 32),
                                                                      d_offset, out);
    counter_add_kernel<<<1, 1, 0, stream>>>(d_offset, (unsigned long long)n);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_init_training_rays(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
                                     uint32_t n_images, uint32_t n_pixels_per_image, int32_t img_w,
                                     int32_t img_h, float n_rays_per_image, const tn_bbox_t* bbox,
                                     const tn_camera_t* cameras, const uint8_t* images, const float* random,
                                     float* pix_rgba, uint32_t* ray_img_id, float* ori, float* dir,
                                     float* ray_t, float* ray_tmax, uint8_t* ray_alive, hipStream_t stream) {
    TN_CHECK_ARG(bbox && cameras && images && random && pix_rgba && ray_img_id && ori && dir && ray_t &&
                 ray_tmax && ray_alive);
    TN_CHECK_ARG(n_rays_max <= stride && n_images > 0 && img_w > 0 && img_h > 0);
    TN_CHECK_ARG(n_pixels_per_image == (uint32_t)img_w * (uint32_t)img_h);
    if (n_rays_max == 0) return TN_OK;
    init_training_rays_kernel<<<blocks_for(n_rays_max, 256), 256, 0, stream>>>(
        n_rays_max, d_n_rays, stride, n_images, n_pixels_per_image, img_w, n_rays_per_image, bbox, cameras,
        images, random, pix_rgba, ray_img_id, ori, dir, ray_t, ray_tmax, ray_alive);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_deactivate_rays(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
                                  float alpha_threshold, float selection_probability, const float* pix_rgba,
                                  const float* random, uint8_t* ray_alive, hipStream_t stream) {
    TN_CHECK_ARG(pix_rgba && random && ray_alive && n_rays_max <= stride);
    if (n_rays_max == 0) return TN_OK;
    deactivate_rays_kernel<<<blocks_for(n_rays_max, 256), 256, 0, stream>>>(
        n_rays_max, d_n_rays, stride, alpha_threshold, selection_probability, pix_rgba, random, ray_alive);
    TN_CHECK_LAUNCH();
    return TN_OK;
}
