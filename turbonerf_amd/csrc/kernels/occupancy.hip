// Occupancy-grid maintenance (reference src/utils/occupancy-grid-kernels.cu,
// src/utils/bit-utils.*, src/core/trainer.cu:221-311) and the analytic
// Gaussian-blob field used for synthetic scenes.
#pragma clang fp contract(off)

#include "common.hpp"

namespace tn {

// density *= factor for every cascade (reference :16-34); float4 streaming
__global__ void occ_decay_kernel(uint64_t n, float factor, float* __restrict__ density) {
    const uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 3 < n) {
        float4 v = *reinterpret_cast<float4*>(density + i);
        v.x *= factor;
        v.y *= factor;
        v.z *= factor;
        v.w *= factor;
        *reinterpret_cast<float4*>(density + i) = v;
    } else {
        for (uint64_t j = i; j < n; ++j) density[j] *= factor;
    }
}

// jittered point in cell (start + i) of `level`, normalised by the training
// box, not clamped (reference :38-76, occupancy-grid.cuh:167-182)
__global__ void occ_sample_points_kernel(uint32_t n_cells, uint32_t B, uint32_t start_idx, int res, int level,
                                         float inv_aabb, const float* __restrict__ random, float* __restrict__ pos) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cells) return;
    const uint32_t idx = i + start_idx;
    const float inv_res = 1.0f / (float)res;
    const int half_res = res / 2;
    const float level_size = (float)(1 << level);
    const float voxel = level_size * inv_res;
    const int ix = (int)compact_bits10(idx >> 0);
    const int iy = (int)compact_bits10(idx >> 1);
    const int iz = (int)compact_bits10(idx >> 2);
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

// sigma_cell = max(sigma_cell, exp(min(s,12)-1)), with the reference's random
// subset rule once training is past warm-up (reference :79-115)
template <typename TIn>
__global__ void occ_update_density_kernel(uint32_t n_cells, uint32_t start_idx, uint32_t volume, int level,
                                          int sample_all, const float* __restrict__ random,
                                          const TIn* __restrict__ raw, const uint8_t* __restrict__ bits,
                                          float* __restrict__ density) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cells) return;
    const uint32_t idx = i + start_idx;
    if (!sample_all) {
        const float r = random[i];
        if (r > 0.5f) return;
        if (!(r < 0.25f) && !((bits[idx] >> level) & 1u)) return;
    }
    const float sigma = __expf(fminf((float)raw[i], 12.0f) - 1.0f);
    float* p = density + (uint64_t)level * volume + idx;
    *p = fmaxf(sigma, *p);
}

// byte = sum_k [density_k > threshold] << k (reference :118-141); one thread
// per 4 cells so the byte stores are one 32-bit store
__global__ void occ_bits_kernel(uint32_t volume, int n_levels, float thr, const float* __restrict__ density,
                                uint8_t* __restrict__ bits) {
    const uint32_t i4 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i4 >= volume) return;
    if (i4 + 3 < volume) {
        uint32_t packed = 0;
        for (int l = 0; l < n_levels; ++l) {
            const float4 d = *reinterpret_cast<const float4*>(density + (uint64_t)l * volume + i4);
            packed |= (uint32_t)(d.x > thr) << l;
            packed |= (uint32_t)(d.y > thr) << (l + 8);
            packed |= (uint32_t)(d.z > thr) << (l + 16);
            packed |= (uint32_t)(d.w > thr) << (l + 24);
        }
        *reinterpret_cast<uint32_t*>(bits + i4) = packed;
    } else {
        for (uint32_t i = i4; i < volume; ++i) {
            uint8_t b = 0;
            for (int l = 0; l < n_levels; ++l) b |= (uint8_t)((density[(uint64_t)l * volume + i] > thr) << l);
            bits[i] = b;
        }
    }
}

// population count of a byte array into out[0] (u64 via 32-bit atomics on
// two words is avoided: one workgroup per chunk writes a partial, a second
// launch sums the partials)
constexpr int kPopThreads = 256;
__global__ void popcount_partial_kernel(uint64_t n, const uint8_t* __restrict__ bytes, uint32_t* __restrict__ part) {
    __shared__ uint32_t sm[kPopThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kPopThreads * 16;
    uint32_t c = 0;
    for (int k = 0; k < 4; ++k) {
        const uint64_t i = base + ((uint64_t)k * kPopThreads + threadIdx.x) * 4;
        if (i + 3 < n) c += __popc(*reinterpret_cast<const uint32_t*>(bytes + i));
        else
            for (uint64_t j = i; j < n; ++j) c += __popc((uint32_t)bytes[j]);
    }
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_down(c, d, 64);
    if (__lane_id() == 0) sm[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0] + sm[1] + sm[2] + sm[3];
}

__global__ void popcount_final_kernel(uint32_t n_part, const uint32_t* __restrict__ part, uint32_t* __restrict__ out) {
    __shared__ uint32_t sm[16];
    uint32_t c = 0;
    for (uint32_t i = threadIdx.x; i < n_part; i += blockDim.x) c += part[i];
    for (int d = 32; d >= 1; d >>= 1) c += __shfl_down(c, d, 64);
    if (__lane_id() == 0) sm[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) t += sm[k];
        out[0] = t;
    }
}

// ---------------------------------------------------------------------------
// analytic blob field: blob k = {cx, cy, cz, s, p, r, g, b} (world units)
// ---------------------------------------------------------------------------
constexpr int kMaxBlobs = 256;

TN_DEVICE void blob_field(const float* __restrict__ bl, uint32_t nb, float x, float y, float z, float& sigma,
                          float& r, float& g, float& b) {
    float s = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
    for (uint32_t k = 0; k < nb; ++k) {
        const float* q = bl + 8 * k;
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float d2 = dx * dx + dy * dy + dz * dz;
        const float rho = q[4] * __expf(-d2 / (2.0f * q[3] * q[3]));
        s += rho;
        sr += rho * q[5];
        sg += rho * q[6];
        sb += rho * q[7];
    }
    sigma = s;
    const float inv = s > 0.0f ? 1.0f / s : 0.0f;
    r = sr * inv;
    g = sg * inv;
    b = sb * inv;
}

// density of the blob field at every cell centre of every cascade
__global__ void blob_density_grid_kernel(uint32_t volume, int n_levels, int res, const float* __restrict__ blobs,
                                         uint32_t nb, float* __restrict__ density) {
    __shared__ float sb[kMaxBlobs * 8];
    for (uint32_t k = threadIdx.x; k < nb * 8; k += blockDim.x) sb[k] = blobs[k];
    __syncthreads();
    const uint64_t gi = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= (uint64_t)volume * n_levels) return;
    const int level = (int)(gi / volume);
    const uint32_t idx = (uint32_t)(gi % volume);
    const float cell = (float)(1 << level) / (float)res;
    const float x = ((float)(int)compact_bits10(idx >> 0) - (float)(res / 2) + 0.5f) * cell;
    const float y = ((float)(int)compact_bits10(idx >> 1) - (float)(res / 2) + 0.5f) * cell;
    const float z = ((float)(int)compact_bits10(idx >> 2) - (float)(res / 2) + 0.5f) * cell;
    float sigma, r, g, b;
    blob_field(sb, nb, x, y, z, sigma, r, g, b);
    density[gi] = sigma;
}

// sigma and rgb of the blob field at sample positions given in normalised
// [0,1] box coordinates (world = (p - 0.5) * aabb)
__global__ void blob_eval_kernel(const uint32_t* __restrict__ info, uint32_t n_max, uint32_t B, float aabb,
                                 const float* __restrict__ blobs, uint32_t nb, const float* __restrict__ pos,
                                 float* __restrict__ sigma, float* __restrict__ rgb) {
    __shared__ float sb[kMaxBlobs * 8];
    for (uint32_t k = threadIdx.x; k < nb * 8; k += blockDim.x) sb[k] = blobs[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    if (i >= n) return;
    const float x = (pos[i] - 0.5f) * aabb, y = (pos[i + B] - 0.5f) * aabb, z = (pos[i + 2 * B] - 0.5f) * aabb;
    float s, r, g, b;
    blob_field(sb, nb, x, y, z, s, r, g, b);
    sigma[i] = s;
    rgb[i] = r;
    rgb[i + B] = g;
    rgb[i + 2 * B] = b;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_occ_decay(tn_occupancy_grid_t grid, float factor, hipStream_t stream) {
    TN_CHECK_ARG(grid.density && grid.n_levels >= 1 && grid.resolution >= 2);
    const uint64_t n = (uint64_t)grid.n_levels * grid.resolution * grid.resolution * grid.resolution;
    occ_decay_kernel<<<blocks_for((n + 3) / 4, 256), 256, 0, stream>>>(n, factor, grid.density);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_occ_sample_points(uint32_t n_cells, uint32_t stride, uint32_t start_idx, tn_occupancy_grid_t grid,
                                    int level, float inv_aabb, const float* random, float* pos, hipStream_t stream) {
    TN_CHECK_ARG(random && pos && n_cells <= stride && level >= 0 && level < grid.n_levels);
    const uint32_t volume = (uint32_t)(grid.resolution * grid.resolution * grid.resolution);
    TN_CHECK_ARG((uint64_t)start_idx + n_cells <= volume);
    if (n_cells == 0) return TN_OK;
    occ_sample_points_kernel<<<blocks_for(n_cells, 256), 256, 0, stream>>>(n_cells, stride, start_idx, grid.resolution,
                                                                           level, inv_aabb, random, pos);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// raw_dtype: 0 = fp32 network output, 1 = fp16
extern "C" int tn_occ_update_density(int raw_dtype, uint32_t n_cells, uint32_t start_idx, tn_occupancy_grid_t grid,
                                     int level, int sample_all_cells, const float* random, const void* raw_density,
                                     hipStream_t stream) {
    TN_CHECK_ARG(random && raw_density && grid.density && grid.bitfield && level >= 0 && level < grid.n_levels);
    const uint32_t volume = (uint32_t)(grid.resolution * grid.resolution * grid.resolution);
    TN_CHECK_ARG((uint64_t)start_idx + n_cells <= volume);
    if (n_cells == 0) return TN_OK;
    if (raw_dtype == 0)
        occ_update_density_kernel<float><<<blocks_for(n_cells, 256), 256, 0, stream>>>(
            n_cells, start_idx, volume, level, sample_all_cells, random, static_cast<const float*>(raw_density),
            grid.bitfield, grid.density);
    else if (raw_dtype == 1)
        occ_update_density_kernel<_Float16><<<blocks_for(n_cells, 256), 256, 0, stream>>>(
            n_cells, start_idx, volume, level, sample_all_cells, random, static_cast<const _Float16*>(raw_density),
            grid.bitfield, grid.density);
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_occ_update_bits(tn_occupancy_grid_t grid, float threshold, hipStream_t stream) {
    TN_CHECK_ARG(grid.density && grid.bitfield && grid.n_levels >= 1 && grid.n_levels <= 8);
    const uint32_t volume = (uint32_t)(grid.resolution * grid.resolution * grid.resolution);
    occ_bits_kernel<<<blocks_for((volume + 3) / 4, 256), 256, 0, stream>>>(volume, grid.n_levels, threshold,
                                                                            grid.density, grid.bitfield);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_popcount_workspace_bytes(uint64_t n_bytes) {
    return (size_t)blocks_for(n_bytes, kPopThreads * 16) * sizeof(uint32_t);
}

// number of set bits in n_bytes bytes -> *out (device u32)
extern "C" int tn_popcount(const uint8_t* bytes, uint64_t n_bytes, uint32_t* out, void* workspace,
                           size_t workspace_bytes, hipStream_t stream) {
    TN_CHECK_ARG(bytes && out && workspace && n_bytes > 0);
    if (workspace_bytes < tn_popcount_workspace_bytes(n_bytes)) return TN_ERR_OUT_OF_MEMORY;
    const unsigned nb = blocks_for(n_bytes, kPopThreads * 16);
    popcount_partial_kernel<<<nb, kPopThreads, 0, stream>>>(n_bytes, bytes, static_cast<uint32_t*>(workspace));
    popcount_final_kernel<<<1, 1024, 0, stream>>>(nb, static_cast<uint32_t*>(workspace), out);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_blob_density_grid(tn_occupancy_grid_t grid, const float* blobs, uint32_t n_blobs,
                                    hipStream_t stream) {
    TN_CHECK_ARG(grid.density && blobs && n_blobs >= 1 && n_blobs <= kMaxBlobs);
    const uint32_t volume = (uint32_t)(grid.resolution * grid.resolution * grid.resolution);
    const uint64_t n = (uint64_t)volume * grid.n_levels;
    blob_density_grid_kernel<<<blocks_for(n, 256), 256, 0, stream>>>(volume, grid.n_levels, grid.resolution, blobs,
                                                                      n_blobs, grid.density);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_blob_eval(const uint32_t* info, uint32_t n_max, uint32_t stride, float aabb, const float* blobs,
                            uint32_t n_blobs, const float* pos, float* sigma, float* rgb, hipStream_t stream) {
    TN_CHECK_ARG(blobs && pos && sigma && rgb && n_blobs >= 1 && n_blobs <= kMaxBlobs && n_max <= stride);
    if (n_max == 0) return TN_OK;
    blob_eval_kernel<<<blocks_for(n_max, 256), 256, 0, stream>>>(info, n_max, stride, aabb, blobs, n_blobs, pos,
                                                                  sigma, rgb);
    TN_CHECK_LAUNCH();
    return TN_OK;
}
