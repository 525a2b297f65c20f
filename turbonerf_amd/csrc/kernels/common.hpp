// Shared device helpers for the TurboNeRF gfx950 kernel library.
//
// Functions used by the bit-exact marching path are written so that, with
// floating-point contraction disabled (`#pragma clang fp contract(off)` in the
// including translation unit), every fp32 operation is a single IEEE
// round-to-nearest operation in the same order as the reference, so integer
// results (sample counts, offsets, Morton indices) match a CPU evaluation of
// the same formulas bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/turbonerf/tn_types.h"

namespace tn {

constexpr int kWave = 64;

#define TN_DEVICE __device__ __forceinline__

// ---------------------------------------------------------------------------
// error handling helpers for the C ABI
// ---------------------------------------------------------------------------
#define TN_CHECK_ARG(cond)                        \
    do {                                          \
        if (!(cond)) return TN_ERR_INVALID_ARGUMENT; \
    } while (0)

#define TN_CHECK_LAUNCH()                                   \
    do {                                                    \
        if (hipGetLastError() != hipSuccess) return TN_ERR_LAUNCH; \
    } while (0)

inline unsigned blocks_for(uint64_t n, unsigned threads) {
    uint64_t b = (n + threads - 1) / threads;
    return (unsigned)(b == 0 ? 1 : b);
}

// Grid of a kernel whose waves loop over the device-side ray count (a wave
// per ray): at most 1024 workgroups, since a training batch uses a small
// part of its ray capacity and idle workgroups still cost their dispatch
inline unsigned ray_wave_blocks(uint64_t n_rays_max, unsigned waves_per_block) {
    const unsigned b = blocks_for(n_rays_max, waves_per_block);
    return b < 1024u ? b : 1024u;
}

// kernels using more than the default 64 KiB of dynamic LDS opt in once
template <typename K>
inline bool allow_lds(K kernel, size_t bytes) {
    if (bytes <= 65536) return true;
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)bytes) == hipSuccess)
        return true;
    (void)hipGetLastError();  // do not leave the failure for the next launch check
    return false;
}

// multiprocessors (CUs) of the current device, cached per device
inline unsigned compute_units() {
    static int cached_dev = -1;
    static unsigned cached_cus = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (dev != cached_dev) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
        cached_cus = (unsigned)cus;
        cached_dev = dev;
    }
    return cached_cus;
}

// Launch-shape knobs (enum tn_tune_knob), defined in tuning.hip; read by the
// launchers on the host when they enqueue a kernel.
int tune_value(int knob);

// ---------------------------------------------------------------------------
// Philox-4x32-10 counter-based RNG (Salmon et al., SC'11). Element g of a
// stream keyed by `seed` is word (g & 3) of philox(counter = g >> 2).
// ---------------------------------------------------------------------------
TN_DEVICE void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
    }
}

TN_DEVICE float u32_to_unit(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------
// Morton code (3-D bit interleave, x -> bit 0)
// ---------------------------------------------------------------------------
TN_DEVICE uint32_t expand_bits10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

TN_DEVICE uint32_t morton3d(uint32_t x, uint32_t y, uint32_t z) {
    return expand_bits10(x) | (expand_bits10(y) << 1) | (expand_bits10(z) << 2);
}

TN_DEVICE uint32_t compact_bits10(uint32_t v) {
    v &= 0x09249249u;
    v = (v | (v >> 2)) & 0x030C30C3u;
    v = (v | (v >> 4)) & 0x0300F00Fu;
    v = (v | (v >> 8)) & 0x030000FFu;
    v = (v | (v >> 16)) & 0x000003FFu;
    return v;
}

// ---------------------------------------------------------------------------
// Directed-rounding division a / b rounded toward +infinity. q = RN(a/b) is
// correctly rounded (HIP's default fp32 division), and the residual
// a - q*b is exact when formed with one fused multiply-add (Boldo & Muller,
// "Exact and approximated error of the FMA", IEEE TC 2011). A residual with
// the sign of b means the exact quotient lies above q.
// ---------------------------------------------------------------------------
TN_DEVICE float fdiv_ru(float a, float b) {
    float q = a / b;
    const float r = __builtin_fmaf(-q, b, a);
    const bool above = (b > 0.0f) ? (r > 0.0f) : (r < 0.0f);
    if (above) {
        uint32_t u = __float_as_uint(q);
        if (q > 0.0f) u += 1u;
        else if (q < 0.0f) u -= 1u;
        else u = 1u;  // +0 -> smallest positive subnormal
        q = __uint_as_float(u);
    }
    return q;
}

// ---------------------------------------------------------------------------
// bounding box (same semantics as the reference BoundingBox)
// ---------------------------------------------------------------------------
TN_DEVICE bool bbox_contains(const tn_bbox_t& b, float x, float y, float z) {
    return x >= b.min_x && x <= b.max_x && y >= b.min_y && y <= b.max_y && z >= b.min_z && z <= b.max_z;
}

TN_DEVICE bool bbox_intersect(const tn_bbox_t& b, float ox, float oy, float oz, float idx, float idy,
                              float idz, float& tmin, float& tmax) {
    float t1x = (b.min_x - ox) * idx, t2x = (b.max_x - ox) * idx;
    float t1y = (b.min_y - oy) * idy, t2y = (b.max_y - oy) * idy;
    float t1z = (b.min_z - oz) * idz, t2z = (b.max_z - oz) * idz;
    if (t1x > t2x) { float s = t1x; t1x = t2x; t2x = s; }
    if (t1y > t2y) { float s = t1y; t1y = t2y; t2y = s; }
    if (t1z > t2z) { float s = t1z; t1z = t2z; t2z = s; }
    tmin = fmaxf(t1x, fmaxf(t1y, t1z));
    tmax = fminf(t2x, fminf(t2y, t2z));
    return tmax >= tmin;
}

// 2^ceil(log2(n)) for n >= 1 (1 for n <= 1). Training gradients are scaled
// by it (and unscaled by its reciprocal, both exact) so that the per-ray mean
// of the loss does not push fp16 gradients towards the subnormal range.
TN_DEVICE float ray_pow2(uint32_t n) { return n <= 1u ? 1.0f : ldexpf(1.0f, 32 - __clz(n - 1u)); }

// ---------------------------------------------------------------------------
// occupancy grid queries
// ---------------------------------------------------------------------------
TN_DEVICE float get_dt(float t, float cone, float dt_min, float dt_max) {
    return fminf(fmaxf(t * cone, dt_min), dt_max);
}

// finest cascade covering p whose cell side exceeds dt, else the coarsest
TN_DEVICE int grid_level_at(int n_levels, float inv_res, float x, float y, float z, float dt) {
    for (int k = 0; k < n_levels; ++k) {
        const float level_size = (float)(1 << k);
        const float half = 0.5f * level_size;
        const float cell = level_size * inv_res;
        if (fabsf(x) < half && fabsf(y) < half && fabsf(z) < half && cell > dt) return k;
    }
    return n_levels - 1;
}

// Morton index of the cell holding p in cascade `level`; each axis index is
// clamped to [0, res-1] (a point exactly on the +extent face maps to the last
// cell instead of one past it).
TN_DEVICE uint32_t voxel_morton_index(float res_f, uint32_t lim, int level, float x, float y, float z) {
    const float s = 1.0f / (float)(1 << level);
    uint32_t xu = (uint32_t)fmaxf((x * s + 0.5f) * res_f, 0.0f);
    uint32_t yu = (uint32_t)fmaxf((y * s + 0.5f) * res_f, 0.0f);
    uint32_t zu = (uint32_t)fmaxf((z * s + 0.5f) * res_f, 0.0f);
    xu = min(xu, lim);
    yu = min(yu, lim);
    zu = min(zu, lim);
    return morton3d(xu, yu, zu);
}

// distance along the ray to the next cell boundary of `level`, rounded up to
// a whole multiple of dt_min through a directed division
TN_DEVICE float dt_to_next_voxel(float res_f, float inv_res, float px, float py, float pz, float dx,
                                 float dy, float dz, float idx, float idy, float idz, float dt_min,
                                 int level) {
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
    return fdiv_ru(t_target, dt_min) * dt_min;
}

// sigma = exp(min(s, 12) - 1) (reference common-network-kernels.cuh:7-10)
TN_DEVICE float density_to_sigma(float s) { return __expf(fminf(s, 12.0f) - 1.0f); }

// ---------------------------------------------------------------------------
// wave-level primitives (wave64)
// ---------------------------------------------------------------------------
template <typename T>
TN_DEVICE T wave_shfl_up(T v, int d) {
    return __shfl_up(v, d, kWave);
}

TN_DEVICE uint32_t lane_id() { return __lane_id(); }

// ---------------------------------------------------------------------------
// NGP Adam on one parameter (reference src/core/adam-optimizer.cuh:48-96).
// ---------------------------------------------------------------------------
// bias-corrected step-size factor sqrt(1 - beta2^t) / (1 - beta1^t)
TN_DEVICE float adam_bias(uint32_t t, float beta1, float beta2) {
    return sqrtf(1.0f - powf(beta2, (float)t)) / (1.0f - powf(beta1, (float)t));
}

// Learning rate of the step: lr0 / (1 + decay * step) with the step read on
// the device when d_step is given (reference schedule, nerf-network.cu:256),
// else lr0 as passed. Contraction off: the same float expression as the host's.
TN_DEVICE float step_lr(float lr0, const uint32_t* d_step, float decay) {
#pragma clang fp contract(off)
    return d_step ? lr0 / (1.0f + decay * (float)(*d_step)) : lr0;
}

// moments and weight after gradient g with bias factor c (no contraction, so
// every kernel that inlines it rounds identically)
TN_DEVICE void adam_apply(float g, float c, float lr, float beta1, float beta2, float eps, float& w, float& m,
                          float& v) {
#pragma clang fp contract(off)
    const float m1 = beta1 * m + (1.0f - beta1) * g;
    const float m2 = beta2 * v + (1.0f - beta2) * (g * g);
    w = w - (lr * c) / (sqrtf(m2) + eps) * m1;
    m = m1;
    v = m2;
}

}  // namespace tn
