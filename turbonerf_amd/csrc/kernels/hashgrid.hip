// Multiresolution hash-grid encoding (Instant-NGP, Mueller et al. 2022,
// sec. 3), the density network's input encoding configured at reference
// src/core/nerf-network.cu:86-104 (L = 16, F = 2, T = 2^19, N_min = 16,
// b = exp((ln N_max - ln N_min)/(L-1)), trilinear).
//
// Layout (tn_hashgrid_t): level l has scale s_l = N_min * b^l - 1 and
// resolution ceil(s_l) + 1; it holds min(roundup8(res^3), T) entries of F
// features, indexed densely (x + y*res + z*res^2) when res^3 fits, otherwise
// by the XOR-of-primes spatial hash (1, 2654435761, 805459861), both taken
// modulo the level size. The table is entry-major: entry e's F features are
// contiguous. Features are written feature-major, out[(l*F+f)*stride + i].
//
// Launch shape. Forward: one thread per (sample, level), blockIdx.y = level,
// so the blocks in flight gather from one level's table at a time. Backward:
// two threads per (sample, level), one per feature, in a 1-D grid whose
// block b serves level b % L and sample chunk b / L; workgroups go to the 8
// XCDs round-robin, so with L = 16 each XCD scatters into two levels'
// gradients only, and the blocks in flight cover every level, which spreads
// the atomics over the whole table. Consecutive
// samples of a ray share coarse-level cells; the backward pass merges runs of
// equal corner indices across the wave with a segmented sum before issuing
// its atomics.
// Cell selection (x*scale + 0.5, floor) is kept free of contraction so it
// matches a CPU evaluation of the same expression exactly.
#pragma clang fp contract(off)

#include "segscan.hpp"

namespace tn {

struct HashCorners {
    uint32_t idx[8];
    float w[8];
};

TN_DEVICE void hash_corners(const tn_hashgrid_t& g, int l, float px, float py, float pz, HashCorners& hc) {
    const float s = g.scale[l];
    const float x = px * s + 0.5f, y = py * s + 0.5f, z = pz * s + 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    const uint32_t gx = (uint32_t)(int32_t)fx0, gy = (uint32_t)(int32_t)fy0, gz = (uint32_t)(int32_t)fz0;
    const uint32_t size = g.level_size[l];
    const uint32_t off = g.level_offset[l];
    const bool dense = g.level_dense[l] != 0;
    const uint32_t r = g.resolution[l];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const uint32_t cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
        const uint32_t X = gx + cx, Y = gy + cy, Z = gz + cz;
        uint32_t id = dense ? (X + Y * r + Z * r * r) : ((X * 1u) ^ (Y * 2654435761u) ^ (Z * 805459861u));
        // level sizes are multiples of 8; hashed levels are powers of two
        id = dense ? id % size : (id & (size - 1u));
        hc.idx[c] = off + id;
        hc.w[c] = (cx ? fx : 1.0f - fx) * (cy ? fy : 1.0f - fy) * (cz ? fz : 1.0f - fz);
    }
}

template <typename T>
struct Pair;
template <>
struct Pair<float> {
    using V = float2;
    TN_DEVICE static float2 ld(const float* t, uint32_t e) { return reinterpret_cast<const float2*>(t)[e]; }
    TN_DEVICE static float a(float2 v) { return v.x; }
    TN_DEVICE static float b(float2 v) { return v.y; }
};
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
template <>
struct Pair<_Float16> {
    using V = half2v;
    TN_DEVICE static half2v ld(const _Float16* t, uint32_t e) { return reinterpret_cast<const half2v*>(t)[e]; }
    TN_DEVICE static float a(half2v v) { return (float)v.x; }
    TN_DEVICE static float b(half2v v) { return (float)v.y; }
};

template <typename TT, typename TO>
__global__ void __launch_bounds__(256) hashgrid_fwd_kernel(uint32_t n_max, const uint32_t* __restrict__ info,
                                                           uint32_t B, uint32_t OB, tn_hashgrid_t g,
                                                           const float* __restrict__ pos, const TT* __restrict__ table,
                                                           TO* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = blockIdx.y;
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    if (i >= n) return;
    HashCorners hc;
    hash_corners(g, l, pos[i], pos[i + B], pos[i + 2 * (uint64_t)B], hc);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const auto v = Pair<TT>::ld(table, hc.idx[c]);
        f0 += hc.w[c] * Pair<TT>::a(v);
        f1 += hc.w[c] * Pair<TT>::b(v);
    }
    out[(uint64_t)(2 * l) * OB + i] = (TO)f0;
    out[(uint64_t)(2 * l + 1) * OB + i] = (TO)f1;
}

// dtable[e] += sum over samples of w * dL/dfeature. Two lanes per sample,
// one per feature, so each atomic instruction updates both features of an
// entry in one 8-byte span (half the memory-side requests of one lane per
// sample). A wave covers 32 consecutive samples; runs of consecutive samples
// that hit the same entry are merged with a segmented sum first, and a wave
// whose samples all differ skips the merge.
template <typename TG>
__global__ void __launch_bounds__(256) hashgrid_bwd_kernel(uint32_t n_max, const uint32_t* __restrict__ info,
                                                           uint32_t B, uint32_t GB, tn_hashgrid_t g,
                                                           const float* __restrict__ pos, const TG* __restrict__ dfeat,
                                                           float* __restrict__ dtable) {
    const uint32_t L = (uint32_t)g.n_levels;
    const int l = (int)(blockIdx.x % L);
    const uint32_t tid = (blockIdx.x / L) * blockDim.x + threadIdx.x;
    const uint32_t i = tid >> 1;
    const uint32_t f = tid & 1u;
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    // whole waves past the end leave; partial waves keep every lane for the shuffles
    if ((i & ~31u) >= n) return;
    const bool valid = i < n;
    const uint32_t lane = __lane_id();
    HashCorners hc;
    float gf = 0.0f;
    if (valid) {
        hash_corners(g, l, pos[i], pos[i + B], pos[i + 2 * (uint64_t)B], hc);
        gf = (float)dfeat[(uint64_t)(2 * l + f) * GB + i];
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            hc.idx[c] = 0;
            hc.w[c] = 0.0f;
        }
    }
    const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        // runs are maximal stretches of consecutive samples with one entry;
        // the same entry may recur further on (hash collisions), so runs are
        // keyed by their ordinal, not by the entry
        const int ent = valid ? (int)hc.idx[c] : -1 - (int)(lane >> 1);
        const int prev = __shfl_up(ent, 2, 64);
        const bool head = lane < 2 || prev != ent;
        const uint64_t hm = __ballot(head);
        float a = hc.w[c] * gf;
        if (hm != ~0ull) {
            const int key = __builtin_popcountll(hm & le & 0x5555555555555555ull);
#pragma unroll
            for (int d = 2; d < 64; d <<= 1) {
                const float o = __shfl_up(a, d, 64);
                const int ok = __shfl_up(key, d, 64);
                if ((int)lane >= d && ok == key) a += o;
            }
            const int next = __shfl_down(ent, 2, 64);
            const bool tail = valid && (lane >= 62 || next != ent);
            if (tail && a != 0.0f) atomicAdd(dtable + 2 * (uint64_t)hc.idx[c] + f, a);
        } else if (valid && a != 0.0f) {
            atomicAdd(dtable + 2 * (uint64_t)hc.idx[c] + f, a);
        }
    }
}

}  // namespace tn

using namespace tn;

// Fills the level table for L levels of F features, 2^log2_T entries per
// level at most, base resolution N_min and per-level scale b.
extern "C" int tn_hashgrid_layout(int n_levels, int n_features, int log2_hashmap_size, int base_resolution,
                                  double per_level_scale, tn_hashgrid_t* out) {
    TN_CHECK_ARG(out && n_levels >= 1 && n_levels <= TN_HASHGRID_MAX_LEVELS && n_features == 2);
    TN_CHECK_ARG(log2_hashmap_size >= 3 && log2_hashmap_size <= 24 && base_resolution >= 2 && per_level_scale >= 1.0);
    tn_hashgrid_t g = {};
    g.n_levels = n_levels;
    g.n_features = n_features;
    g.log2_hashmap_size = log2_hashmap_size;
    g.base_resolution = base_resolution;
    g.per_level_scale = (float)per_level_scale;
    const uint64_t T = 1ull << log2_hashmap_size;
    uint64_t off = 0;
    for (int l = 0; l < n_levels; ++l) {
        const double s = exp2((double)l * log2(per_level_scale)) * (double)base_resolution - 1.0;
        const float sf = (float)s;
        const uint32_t res = (uint32_t)ceil((double)sf) + 1u;
        const uint64_t vol = (uint64_t)res * res * res;
        uint64_t size = (vol + 7) / 8 * 8;
        if (size > T) size = T;
        g.scale[l] = sf;
        g.resolution[l] = res;
        g.level_size[l] = (uint32_t)size;
        g.level_offset[l] = (uint32_t)off;
        g.level_dense[l] = vol <= size ? 1u : 0u;
        off += size;
    }
    if (off * (uint64_t)n_features >= (1ull << 32)) return TN_ERR_UNSUPPORTED;
    g.n_params = (uint32_t)(off * (uint64_t)n_features);
    *out = g;
    return TN_OK;
}

// table_dtype / feat_dtype: 0 = fp32, 1 = fp16. Features written with stride out_stride.
extern "C" int tn_hashgrid_forward(int table_dtype, int feat_dtype, uint32_t n_max, const uint32_t* info,
                                   uint32_t stride, uint32_t out_stride, tn_hashgrid_t grid, const float* pos,
                                   const void* table, void* features, hipStream_t stream) {
    TN_CHECK_ARG(pos && table && features && n_max <= stride && n_max <= out_stride && grid.n_features == 2);
    TN_CHECK_ARG(grid.n_levels >= 1 && grid.n_levels <= TN_HASHGRID_MAX_LEVELS);
    if (n_max == 0) return TN_OK;
    const dim3 blocks(blocks_for(n_max, 256), grid.n_levels);
    if (table_dtype == 0 && feat_dtype == 0)
        hashgrid_fwd_kernel<float, float><<<blocks, 256, 0, stream>>>(n_max, info, stride, out_stride, grid, pos,
                                                                      static_cast<const float*>(table),
                                                                      static_cast<float*>(features));
    else if (table_dtype == 1 && feat_dtype == 1)
        hashgrid_fwd_kernel<_Float16, _Float16><<<blocks, 256, 0, stream>>>(
            n_max, info, stride, out_stride, grid, pos, static_cast<const _Float16*>(table),
            static_cast<_Float16*>(features));
    else if (table_dtype == 1 && feat_dtype == 0)
        hashgrid_fwd_kernel<_Float16, float><<<blocks, 256, 0, stream>>>(n_max, info, stride, out_stride, grid, pos,
                                                                         static_cast<const _Float16*>(table),
                                                                         static_cast<float*>(features));
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// dfeat_dtype: 0 = fp32, 1 = fp16 feature gradients (stride dfeat_stride);
// the fp32 table gradient is accumulated in place.
extern "C" int tn_hashgrid_backward(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
                                    uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos, const void* dfeat,
                                    float* dtable, hipStream_t stream) {
    TN_CHECK_ARG(pos && dfeat && dtable && n_max <= stride && n_max <= dfeat_stride && grid.n_features == 2);
    TN_CHECK_ARG(grid.n_levels >= 1 && grid.n_levels <= TN_HASHGRID_MAX_LEVELS);
    if (n_max == 0) return TN_OK;
    const unsigned blocks = blocks_for(2 * (uint64_t)n_max, 256) * (unsigned)grid.n_levels;
    if (dfeat_dtype == 0)
        hashgrid_bwd_kernel<float><<<blocks, 256, 0, stream>>>(n_max, info, stride, dfeat_stride, grid, pos,
                                                               static_cast<const float*>(dfeat), dtable);
    else if (dfeat_dtype == 1)
        hashgrid_bwd_kernel<_Float16><<<blocks, 256, 0, stream>>>(n_max, info, stride, dfeat_stride, grid, pos,
                                                                  static_cast<const _Float16*>(dfeat), dtable);
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}
