// Occupancy-grid ray marching and compacted sample generation
// (reference src/utils/training-batch-kernels.cuh:212-481,
//  src/core/occupancy-grid.cuh:150-282, src/core/trainer.cu:102-216).
//
// The count and generate passes run the same DDA with contraction disabled,
// so the samples written by the generate pass are exactly the ones counted.
// The exclusive scan and the prefix selection run on the device and leave
// (n_rays, n_samples) in a device-side tn_batch_info, so a training step has
// no host round trip between marching and the network.
#pragma clang fp contract(off)

#include "common.hpp"

namespace tn {

struct MarchGrid {
    const uint8_t* bits;
    int n_levels;
    float res_f;
    float inv_res;
    uint32_t lim;
    float level_scale;  // the cascade is chosen from dt * level_scale
};

TN_DEVICE MarchGrid make_grid(const tn_occupancy_grid_t& g) {
    MarchGrid m;
    m.bits = g.bitfield;
    m.n_levels = g.n_levels;
    m.res_f = (float)g.resolution;
    m.inv_res = 1.0f / m.res_f;
    m.lim = (uint32_t)g.resolution - 1u;
    m.level_scale = g.level_dt_scale > 0.0f ? g.level_dt_scale : 1.0f;
    return m;
}

// Marching runs one wave per ray and reproduces the reference's serial
// recurrence exactly (contraction off, same operation order). The 64 lanes
// buy two things:
//  * runs of occupied steps: from a point t inside the box, lane j forms t_j,
//    the t after j more occupied steps (the same serial sums), and reads the
//    occupancy bit of its own cell at its own cascade; the leading lanes whose
//    steps are inside, within the limits and occupied are taken at once, for
//    one gather per 64 steps. If lane 0 fails, the cell at t is empty;
//  * empty-space skipping: the serial DDA skip steps look cells up in a wave
//    cache. A lookup that misses refills it with 64 parallel loads, lane k
//    taking the cell at k half-cells further along the ray from the point that
//    missed (t0). Entries are keyed by (cascade, Morton index) and hold that
//    cell's own bit, so a prefetched entry can never change a result; lane 0
//    holds the exact cell of the point that missed.
struct WaveCache {
    uint32_t key = 0xFFFFFFFFu;  // per lane
    uint32_t occ = 0;            // per lane
    int level = -1;              // wave-uniform from here on
    float t0 = 0.0f;             // ray t of lane 0's point
    float inv_h = 0.0f;          // 1 / (half a cell of `level`, in ray t)
};

TN_DEVICE void wave_cache_fill(const MarchGrid& g, int level, float t, float x, float y, float z, float dx, float dy,
                               float dz, WaveCache& wc) {
    const float half_cell = 0.5f * (float)(1 << level) * g.inv_res;
    const float tj = (float)__lane_id() * half_cell;  // 0 for lane 0: its point is (x, y, z) exactly
    const uint32_t m = voxel_morton_index(g.res_f, g.lim, level, x + tj * dx, y + tj * dy, z + tj * dz);
    wc.key = m;
    wc.occ = (g.bits[m] >> level) & 1u;
    wc.level = level;
    wc.t0 = t;
    wc.inv_h = 1.0f / half_cell;
}

TN_DEVICE bool wave_occupied(const MarchGrid& g, int level, float t, float x, float y, float z, float dx, float dy,
                             float dz, WaveCache& wc) {
    const uint32_t m = voxel_morton_index(g.res_f, g.lim, level, x, y, z);
    uint64_t hit = level == wc.level ? __ballot(wc.key == m) : 0ull;
    if (hit == 0ull) {
        wave_cache_fill(g, level, t, x, y, z, dx, dy, dz, wc);
        hit = __ballot(wc.key == m);  // lane 0 at least
    }
    return (__ballot(wc.occ != 0u) & hit) != 0ull;
}

struct MarchRay {
    float ox, oy, oz, dx, dy, dz;
};

// Leading lanes (at least 0, at most 64) whose step, j occupied steps after
// t, is inside the box, below t_max, within `budget` and occupied; t_lane =
// t_j. Each lane still in the running reads its own cell's bit. `empty_stop`
// tells whether the first lane that failed did so only for an empty cell
// (the serial loop's next step is then a DDA skip from that lane's t).
TN_DEVICE uint32_t occupied_run(const MarchGrid& g, const tn_bbox_t& bbox, const MarchRay& r, float t, float t_max,
                                uint32_t budget, float cone, float dt_min, float dt_max, float& t_lane,
                                bool& empty_stop) {
    const uint32_t lane = __lane_id();
    float tj = t;
    if (dt_min == dt_max) {
        for (uint32_t k = 0; k < 63; ++k) {
            const float tn = tj + dt_min;
            tj = lane > k ? tn : tj;
        }
    } else {
        for (uint32_t k = 0; k < 63; ++k) {
            const float tn = tj + get_dt(tj, cone, dt_min, dt_max);
            tj = lane > k ? tn : tj;
        }
    }
    t_lane = tj;
    const float x = r.ox + tj * r.dx, y = r.oy + tj * r.dy, z = r.oz + tj * r.dz;
    const bool within = lane < budget && tj < t_max && bbox_contains(bbox, x, y, z);
    bool good = within;
    if (within) {
        const float dtj = get_dt(tj, cone, dt_min, dt_max);
        const int lvl = grid_level_at(g.n_levels, g.inv_res, x, y, z, dtj * g.level_scale);
        const uint32_t m = voxel_morton_index(g.res_f, g.lim, lvl, x, y, z);
        good = ((g.bits[m] >> lvl) & 1u) != 0u;
    }
    const uint64_t bad = __ballot(!good);
    const uint64_t empty = __ballot(within && !good);
    const uint32_t f = bad == 0ull ? 64u : (uint32_t)__builtin_ctzll(bad);
    empty_stop = f < 64u && ((empty >> f) & 1ull) != 0ull;
    return f;
}

// DDA skip steps from t, whose cell is empty, until the point is outside the
// box, at or beyond t_max, or in an occupied cell (the reference's loop takes
// the same steps, testing each point in turn); returns that point's t.
TN_DEVICE float skip_empty(const MarchGrid& g, const tn_bbox_t& bbox, const MarchRay& r, float t, float t_max,
                           float cone, float dt_min, float dt_max, float idx, float idy, float idz, WaveCache& wc) {
    float x = r.ox + t * r.dx, y = r.oy + t * r.dy, z = r.oz + t * r.dz;
    int lvl = grid_level_at(g.n_levels, g.inv_res, x, y, z, get_dt(t, cone, dt_min, dt_max) * g.level_scale);
    for (;;) {
        t += dt_to_next_voxel(g.res_f, g.inv_res, x, y, z, r.dx, r.dy, r.dz, idx, idy, idz, dt_min, lvl);
        if (!(t < t_max)) return t;
        x = r.ox + t * r.dx;
        y = r.oy + t * r.dy;
        z = r.oz + t * r.dz;
        if (!bbox_contains(bbox, x, y, z)) return t;
        lvl = grid_level_at(g.n_levels, g.inv_res, x, y, z, get_dt(t, cone, dt_min, dt_max) * g.level_scale);
        if (wave_occupied(g, lvl, t, x, y, z, r.dx, r.dy, r.dz, wc)) return t;
    }
}

// t after the run of f steps: lane f's t_f, or one more step after lane 63's
TN_DEVICE float after_run(uint32_t f, float t_lane, float cone, float dt_min, float dt_max) {
    if (f < 64u) return __shfl(t_lane, (int)f, 64);
    const float t63 = __shfl(t_lane, 63, 64);
    return t63 + get_dt(t63, cone, dt_min, dt_max);
}

constexpr int kMarchWaves = 4;  // waves per workgroup
// The marching kernels are launched on at most TN_TUNE_MARCH_MAX_BLOCKS
// workgroups (2048 by default: a larger grid leaves a large scene's render
// passes under four rays per wave, in the wave-per-ray mode: 8192 took its
// 1080p frame from 81 to 199 ms) and stride
// over the live rays, one ray per wave at a time: the ray count lives on the
// device, and a grid sized for the largest batch would launch a wave per
// possible ray, nearly all of which exit at once yet occupy the dispatcher
// beside the network kernels.
inline unsigned march_blocks(uint32_t n_rays_max) {
    return std::min(blocks_for(n_rays_max, kMarchWaves), (unsigned)tune_value(TN_TUNE_MARCH_MAX_BLOCKS));
}

TN_DEVICE void march_count_ray(uint32_t i, uint32_t B, const tn_bbox_t& bbox, const MarchGrid& g, float cone,
                               float dt_min, float dt_max, uint32_t max_steps, const float* __restrict__ dir,
                               const float* __restrict__ ray_tmax, uint8_t* __restrict__ ray_alive,
                               const float* __restrict__ ori, const float* __restrict__ ray_t,
                               uint32_t* __restrict__ n_steps) {
    const bool lane0 = __lane_id() == 0;
    if (!ray_alive[i]) {
        if (lane0) n_steps[i] = 0;
        return;
    }
    const MarchRay r = {ori[i], ori[i + B], ori[i + 2 * B], dir[i], dir[i + B], dir[i + 2 * B]};
    const float idx = 1.0f / r.dx, idy = 1.0f / r.dy, idz = 1.0f / r.dz;
    uint32_t n = 0;
    float t = ray_t[i];
    const float t_max = ray_tmax[i];
    WaveCache wc;
    while (t < t_max && n < max_steps) {
        const float x = r.ox + t * r.dx, y = r.oy + t * r.dy, z = r.oz + t * r.dz;
        if (!bbox_contains(bbox, x, y, z)) break;
        float t_lane;
        bool empty;
        const uint32_t f = occupied_run(g, bbox, r, t, t_max, max_steps - n, cone, dt_min, dt_max, t_lane, empty);
        if (f != 0u) {
            t = after_run(f, t_lane, cone, dt_min, dt_max);
            n += f;
        }
        // f == 0: lane 0 is inside, below t_max and within budget, so its cell is empty
        if (f == 0u || empty) t = skip_empty(g, bbox, r, t, t_max, cone, dt_min, dt_max, idx, idy, idz, wc);
    }
    if (lane0) {
        if (n == 0) ray_alive[i] = 0;
        n_steps[i] = n;
    }
}

TN_DEVICE float lane_f(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }
TN_DEVICE uint32_t lane_u(uint32_t v, int l) { return (uint32_t)__builtin_amdgcn_readlane((int)v, l); }

// Many rays per wave (a render pass or a sparse scene): lanes [0, K) own rays
// first + lane. Each lane takes its ray through empty cells by the serial DDA
// itself (reading its cells' bits), so the long skips of rays through mostly
// empty space run side by side; a ray that reaches an occupied cell is then
// taken by the whole wave for its occupied runs (one ray at a time, as in
// march_count_ray), after which its lane resumes. Every ray follows the same
// recurrence as march_count_ray, step for step.
struct LaneRayState {
    MarchRay r;
    float idx, idy, idz, t, t_max;
    uint32_t n;
    bool done, ready;
};

TN_DEVICE void lane_skip_step(const MarchGrid& g, const tn_bbox_t& bbox, float cone, float dt_min, float dt_max,
                              uint32_t max_steps, LaneRayState& q) {
    if (!(q.t < q.t_max && q.n < max_steps)) {
        q.done = true;
        return;
    }
    const float x = q.r.ox + q.t * q.r.dx, y = q.r.oy + q.t * q.r.dy, z = q.r.oz + q.t * q.r.dz;
    if (!bbox_contains(bbox, x, y, z)) {
        q.done = true;
        return;
    }
    const int lvl = grid_level_at(g.n_levels, g.inv_res, x, y, z, get_dt(q.t, cone, dt_min, dt_max) * g.level_scale);
    if ((g.bits[voxel_morton_index(g.res_f, g.lim, lvl, x, y, z)] >> lvl) & 1u)
        q.ready = true;
    else
        q.t += dt_to_next_voxel(g.res_f, g.inv_res, x, y, z, q.r.dx, q.r.dy, q.r.dz, q.idx, q.idy, q.idz, dt_min, lvl);
}

TN_DEVICE LaneRayState lane_ray_state(bool active, uint32_t i, uint32_t B, const float* __restrict__ ori,
                                      const float* __restrict__ dir, const float* __restrict__ ray_t, float t_max) {
    LaneRayState q;
    q.r = active ? MarchRay{ori[i], ori[i + B], ori[i + 2 * B], dir[i], dir[i + B], dir[i + 2 * B]}
                 : MarchRay{0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f};
    q.idx = 1.0f / q.r.dx;
    q.idy = 1.0f / q.r.dy;
    q.idz = 1.0f / q.r.dz;
    q.t = active ? ray_t[i] : 0.0f;
    q.t_max = t_max;
    q.n = 0;
    q.done = !active;
    q.ready = false;
    return q;
}

TN_DEVICE MarchRay lane_ray(const MarchRay& r, int l) {
    return MarchRay{lane_f(r.ox, l), lane_f(r.oy, l), lane_f(r.oz, l), lane_f(r.dx, l), lane_f(r.dy, l),
                    lane_f(r.dz, l)};
}

TN_DEVICE void march_count_lanes(uint32_t first, uint32_t K, uint32_t n_rays, uint32_t B, const tn_bbox_t& bbox,
                                 const MarchGrid& g, float cone, float dt_min, float dt_max, uint32_t max_steps,
                                 const float* __restrict__ dir, const float* __restrict__ ray_tmax,
                                 uint8_t* __restrict__ ray_alive, const float* __restrict__ ori,
                                 const float* __restrict__ ray_t, uint32_t* __restrict__ n_steps) {
    const uint32_t lane = __lane_id();
    const uint32_t i = first + lane;
    const bool own = lane < K && i < n_rays;
    const bool active = own && ray_alive[i] != 0;
    LaneRayState q = lane_ray_state(active, i, B, ori, dir, ray_t, active ? ray_tmax[i] : 0.0f);
    while (__ballot(!q.done) != 0ull) {
        if (!q.done && !q.ready) lane_skip_step(g, bbox, cone, dt_min, dt_max, max_steps, q);
        uint64_t rb = __ballot(q.ready);
        while (rb != 0ull) {
            const int L = __builtin_ctzll(rb);
            rb &= rb - 1ull;
            const MarchRay r = lane_ray(q.r, L);
            float t = lane_f(q.t, L);
            const float t_max = lane_f(q.t_max, L);
            uint32_t n = lane_u(q.n, L);
            for (;;) {  // the point at t is occupied and within the limits, so the first run takes a step
                float t_lane;
                bool empty;
                const uint32_t f = occupied_run(g, bbox, r, t, t_max, max_steps - n, cone, dt_min, dt_max, t_lane, empty);
                if (f == 0u) break;
                t = after_run(f, t_lane, cone, dt_min, dt_max);
                n += f;
                if (f < 64u || !(t < t_max && n < max_steps)) break;
            }
            if ((int)lane == L) {
                q.t = t;
                q.n = n;
                q.ready = false;  // the lane's next step re-tests the point where the run stopped
            }
        }
    }
    if (own) {
        if (!active) {
            n_steps[i] = 0;
        } else {
            if (q.n == 0) ray_alive[i] = 0;
            n_steps[i] = q.n;
        }
    }
}

// Rays per wave from which lanes take rays through empty space themselves;
// below it a wave marches one ray at a time (march_count_ray), skipping with
// its wave cache, which is faster per ray.
constexpr uint32_t kLaneRaysMin = 4;

TN_DEVICE uint32_t rays_per_wave(uint32_t n_rays, uint32_t n_waves) {
    const uint32_t k = (n_rays + n_waves - 1) / n_waves;
    return k < kLaneRaysMin ? 1u : min(k, 64u);
}

__global__ void __launch_bounds__(64 * kMarchWaves)
march_count_kernel(uint32_t n_rays_max, const uint32_t* __restrict__ d_n_rays, uint32_t B, tn_bbox_t bbox,
                   tn_occupancy_grid_t grid, float cone, float dt_min, float dt_max, uint32_t max_steps,
                   const float* __restrict__ dir, const float* __restrict__ ray_tmax, uint8_t* __restrict__ ray_alive,
                   const float* __restrict__ ori, const float* __restrict__ ray_t, uint32_t* __restrict__ n_steps) {
    const uint32_t n_rays = d_n_rays ? min(*d_n_rays, n_rays_max) : n_rays_max;
    const uint32_t n_waves = gridDim.x * kMarchWaves;
    const uint32_t wave = blockIdx.x * kMarchWaves + (threadIdx.x >> 6);
    const MarchGrid g = make_grid(grid);
    const uint32_t K = rays_per_wave(n_rays, n_waves);
    if (K == 1) {
        for (uint32_t i = wave; i < n_rays; i += n_waves)
            march_count_ray(i, B, bbox, g, cone, dt_min, dt_max, max_steps, dir, ray_tmax, ray_alive, ori, ray_t,
                            n_steps);
    } else {
        for (uint32_t first = wave * K; first < n_rays; first += n_waves * K)
            march_count_lanes(first, K, n_rays, B, bbox, g, cone, dt_min, dt_max, max_steps, dir, ray_tmax, ray_alive,
                              ori, ray_t, n_steps);
    }
}

// ---------------------------------------------------------------------------
// exclusive scan + prefix selection. Three launches: per-tile sums, a single
// workgroup scanning the tile sums (at most kScanTile tiles), and a per-tile
// scan that writes offsets and picks the last ray whose offset is below the
// sample budget.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;  // 4096 rays per tile

// inclusive scan of one value per thread across a 256-thread block
TN_DEVICE uint32_t block_inclusive_scan(uint32_t v, uint32_t* smem /*[kScanThreads/64]*/, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wid = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= (uint32_t)d) v += o;
    }
    if (lane == 63) smem[wid] = v;
    __syncthreads();
    uint32_t pre = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t s = smem[w];
        if ((uint32_t)w < wid) pre += s;
        total += s;
    }
    __syncthreads();
    return v + pre;
}

__global__ void scan_tile_sums_kernel(uint32_t n, const uint32_t* __restrict__ d_n, const uint32_t* __restrict__ vals,
                                      uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t smem[kScanThreads / 64];
    if (d_n) n = min(*d_n, n);
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t j = base + k;
        if (j < n) s += vals[j];
    }
    uint32_t total;
    block_inclusive_scan(s, smem, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void scan_tile_prefix_kernel(uint32_t n_tiles, uint32_t* __restrict__ tile_sums) {
    __shared__ uint32_t smem[kScanThreads / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n_tiles; base += kScanThreads) {
        const uint32_t j = base + threadIdx.x;
        const uint32_t v = j < n_tiles ? tile_sums[j] : 0u;
        uint32_t total;
        const uint32_t inc = block_inclusive_scan(v, smem, total);
        if (j < n_tiles) tile_sums[j] = carry + inc - v;  // exclusive
        carry += total;
    }
}

// info[0] = n_rays kept, info[1] = n_samples, info[2] = status, info[3] =
// rays scanned (the batch's generated ray count).
// mode 0 (training, the reference's rule): j = last ray whose offset is below
//   max_samples; rays [0, j) are kept and n_samples = offset[j].
// mode 1 (rendering): every ray whose samples end within max_samples is kept.
__global__ void scan_write_select_kernel(uint32_t n, const uint32_t* __restrict__ d_n,
                                         const uint32_t* __restrict__ vals,
                                         const uint32_t* __restrict__ tile_prefix, uint32_t max_samples, int mode,
                                         uint32_t* __restrict__ offsets, uint32_t* __restrict__ info) {
    __shared__ uint32_t smem[kScanThreads / 64];
    if (d_n) n = min(*d_n, n);
    if (blockIdx.x == 0 && threadIdx.x == 0) info[3] = n;
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t j = base + k;
        v[k] = j < n ? vals[j] : 0u;
        s += v[k];
    }
    uint32_t total;
    const uint32_t inc = block_inclusive_scan(s, smem, total);
    uint64_t run = (uint64_t)tile_prefix[blockIdx.x] + inc - s;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t j = base + k;
        if (j < n) {
            const uint32_t off = (uint32_t)min(run, (uint64_t)0xFFFFFFFFu);
            offsets[j] = off;
            const uint64_t next = run + v[k];
            // offsets are non-decreasing, so exactly one ray is the last with
            // offset < max_samples; the reference keeps the rays before it.
            if (mode == 0) {
                const bool last_lt = off < max_samples && (j + 1 == n || next >= max_samples);
                if (last_lt) {
                    info[0] = j;
                    info[1] = off;
                    info[2] = j == 0 ? (uint32_t)TN_ERR_NO_RAYS : (off == 0 ? (uint32_t)TN_ERR_NO_SAMPLES : 0u);
                }
            } else {
                // ends are non-decreasing: the last ray ending within budget is unique
                const uint64_t after = (j + 1 < n) ? next + vals[j + 1] : ~0ull;
                const bool last_fit = next <= max_samples && (j + 1 == n || after > max_samples);
                if (last_fit) {
                    info[0] = j + 1;
                    info[1] = (uint32_t)next;
                    info[2] = next == 0 ? (uint32_t)TN_ERR_NO_SAMPLES : 0u;
                }
                if (j == 0 && v[k] > max_samples) {
                    info[0] = 0;
                    info[1] = 0;
                    info[2] = (uint32_t)TN_ERR_NO_RAYS;
                }
            }
            run = next;
        }
    }
}

TN_DEVICE void march_generate_ray(uint32_t i, uint32_t B, const tn_bbox_t& bbox, const MarchGrid& g, float dt_min,
                                  float dt_max, float cone, const float* __restrict__ ori,
                                  const float* __restrict__ dir, const float* __restrict__ ray_t,
                                  const uint32_t* __restrict__ ray_offset, const uint32_t* __restrict__ n_ray_steps,
                                  uint32_t* __restrict__ out_ray, float* __restrict__ out_t) {
    const uint32_t lane = __lane_id();
    const uint32_t n_steps = n_ray_steps[i];
    const uint32_t off = ray_offset[i];
    const MarchRay r = {ori[i], ori[i + B], ori[i + 2 * B], dir[i], dir[i + B], dir[i + 2 * B]};
    const float idx = 1.0f / r.dx, idy = 1.0f / r.dy, idz = 1.0f / r.dz;
    float t = ray_t[i];
    uint32_t n = 0;
    WaveCache wc;
    while (n < n_steps) {
        const float x = r.ox + t * r.dx, y = r.oy + t * r.dy, z = r.oz + t * r.dz;
        if (!bbox_contains(bbox, x, y, z)) {
            // the count pass stops at the box boundary, so every counted
            // sample is inside; one outside would be emitted without advancing
            if (lane == 0) {
                out_t[off + n] = t;
                out_ray[off + n] = i;
            }
            ++n;
            continue;
        }
        float t_lane;
        bool empty;
        const uint32_t f = occupied_run(g, bbox, r, t, INFINITY, n_steps - n, cone, dt_min, dt_max, t_lane, empty);
        if (f != 0u) {
            if (lane < f) {  // one coalesced row of samples
                out_t[off + n + lane] = t_lane;
                out_ray[off + n + lane] = i;
            }
            t = after_run(f, t_lane, cone, dt_min, dt_max);
            n += f;
        }
        if (f == 0u || empty) t = skip_empty(g, bbox, r, t, INFINITY, cone, dt_min, dt_max, idx, idy, idz, wc);
    }
}

// march_generate_ray for rays first + lane, lanes [0, K) (see march_count_lanes)
TN_DEVICE void march_generate_lanes(uint32_t first, uint32_t K, uint32_t n_rays, uint32_t B, const tn_bbox_t& bbox,
                                    const MarchGrid& g, float dt_min, float dt_max, float cone,
                                    const float* __restrict__ ori, const float* __restrict__ dir,
                                    const float* __restrict__ ray_t, const uint32_t* __restrict__ ray_offset,
                                    const uint8_t* __restrict__ ray_alive, const uint32_t* __restrict__ n_ray_steps,
                                    uint32_t* __restrict__ out_ray, float* __restrict__ out_t) {
    const uint32_t lane = __lane_id();
    const uint32_t i = first + lane;
    const bool active = lane < K && i < n_rays && ray_alive[i] != 0;
    const uint32_t ns = active ? n_ray_steps[i] : 0u;
    const uint32_t off = active ? ray_offset[i] : 0u;
    LaneRayState q = lane_ray_state(active, i, B, ori, dir, ray_t, INFINITY);
    while (__ballot(!q.done) != 0ull) {
        if (!q.done && !q.ready) {
            if (!(q.n < ns)) {
                q.done = true;
            } else {
                const float x = q.r.ox + q.t * q.r.dx, y = q.r.oy + q.t * q.r.dy, z = q.r.oz + q.t * q.r.dz;
                if (!bbox_contains(bbox, x, y, z)) {
                    // the count pass stops at the box boundary, so every counted
                    // sample is inside; one outside would be emitted without advancing
                    out_t[off + q.n] = q.t;
                    out_ray[off + q.n] = i;
                    ++q.n;
                } else {
                    const int lvl = grid_level_at(g.n_levels, g.inv_res, x, y, z, get_dt(q.t, cone, dt_min, dt_max) * g.level_scale);
                    if ((g.bits[voxel_morton_index(g.res_f, g.lim, lvl, x, y, z)] >> lvl) & 1u)
                        q.ready = true;
                    else
                        q.t += dt_to_next_voxel(g.res_f, g.inv_res, x, y, z, q.r.dx, q.r.dy, q.r.dz, q.idx, q.idy,
                                                q.idz, dt_min, lvl);
                }
            }
        }
        uint64_t rb = __ballot(q.ready);
        while (rb != 0ull) {
            const int L = __builtin_ctzll(rb);
            rb &= rb - 1ull;
            const MarchRay r = lane_ray(q.r, L);
            float t = lane_f(q.t, L);
            uint32_t n = lane_u(q.n, L);
            const uint32_t nsL = lane_u(ns, L), offL = lane_u(off, L), iL = lane_u(i, L);
            for (;;) {
                float t_lane;
                bool empty;
                const uint32_t f = occupied_run(g, bbox, r, t, INFINITY, nsL - n, cone, dt_min, dt_max, t_lane, empty);
                if (f == 0u) break;
                if (lane < f) {  // one coalesced row of samples
                    out_t[offL + n + lane] = t_lane;
                    out_ray[offL + n + lane] = iL;
                }
                t = after_run(f, t_lane, cone, dt_min, dt_max);
                n += f;
                if (f < 64u || n >= nsL) break;
            }
            if ((int)lane == L) {
                q.t = t;
                q.n = n;
                q.ready = false;
            }
        }
    }
}

__global__ void __launch_bounds__(64 * kMarchWaves)
march_generate_kernel(uint32_t n_rays_max, const uint32_t* __restrict__ info, uint32_t B, tn_bbox_t bbox,
                      tn_occupancy_grid_t grid, float dt_min, float dt_max, float cone, const float* __restrict__ ori,
                      const float* __restrict__ dir, const float* __restrict__ ray_t,
                      const uint32_t* __restrict__ ray_offset, const uint8_t* __restrict__ ray_alive,
                      const uint32_t* __restrict__ n_ray_steps, uint32_t* __restrict__ out_ray,
                      float* __restrict__ out_t) {
    const uint32_t n_rays = min(info[0], n_rays_max);
    const uint32_t n_waves = gridDim.x * kMarchWaves;
    const uint32_t wave = blockIdx.x * kMarchWaves + (threadIdx.x >> 6);
    const MarchGrid g = make_grid(grid);
    const uint32_t K = rays_per_wave(n_rays, n_waves);
    if (K == 1) {
        for (uint32_t i = wave; i < n_rays; i += n_waves) {
            if (ray_alive[i])
                march_generate_ray(i, B, bbox, g, dt_min, dt_max, cone, ori, dir, ray_t, ray_offset, n_ray_steps,
                                   out_ray, out_t);
        }
    } else {
        for (uint32_t first = wave * K; first < n_rays; first += n_waves * K)
            march_generate_lanes(first, K, n_rays, B, bbox, g, dt_min, dt_max, cone, ori, dir, ray_t, ray_offset,
                                 ray_alive, n_ray_steps, out_ray, out_t);
    }
}

// Sample-parallel second half of generation: from each sample's t and ray
// (left in out_img_id by the marching pass) write the network inputs and the
// distortion loss's normalised midpoints and steps. The ray's span ends where
// its last sample's step ends, fl(t_last + dt(t_last)), which is where the
// marching pass's t stood after its last emission.
__global__ void march_expand_kernel(uint32_t n_rays_max, uint32_t SB_max, const uint32_t* __restrict__ info,
                                    uint32_t B, uint32_t SB, float inv_aabb, float dt_min, float dt_max, float cone,
                                    const uint32_t* __restrict__ ray_img_id, const float* __restrict__ ori,
                                    const float* __restrict__ dir, const float* __restrict__ ray_t,
                                    const uint32_t* __restrict__ ray_offset, const uint32_t* __restrict__ n_ray_steps,
                                    uint32_t* __restrict__ out_img_id, const float* __restrict__ out_t,
                                    float* __restrict__ out_pos, float* __restrict__ out_dir, float* __restrict__ out_dt,
                                    float* __restrict__ out_m_norm, float* __restrict__ out_dt_norm) {
    const uint32_t sid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_samples = min(info[1], SB_max);
    if (sid >= n_samples) return;
    const uint32_t r = out_img_id[sid];
    if (r >= n_rays_max) return;  // defensive: every sample below n_samples was written by the marching pass
    const float t = out_t[sid];
    const float dt = get_dt(t, cone, dt_min, dt_max);
    const float x = ori[r] + t * dir[r];
    const float y = ori[r + B] + t * dir[r + B];
    const float z = ori[r + 2 * B] + t * dir[r + 2 * B];
    const float t_min = ray_t[r];
    const float t_last = out_t[ray_offset[r] + n_ray_steps[r] - 1u];
    const float t_max = t_last + get_dt(t_last, cone, dt_min, dt_max);
    const float t_span = fmaxf(t_max - t_min, 1e-6f);
    out_img_id[sid] = ray_img_id[r];
    out_dt[sid] = inv_aabb * dt;
    out_pos[sid] = fminf(fmaxf(x * inv_aabb + 0.5f, 0.0f), 1.0f);
    out_pos[sid + SB] = fminf(fmaxf(y * inv_aabb + 0.5f, 0.0f), 1.0f);
    out_pos[sid + 2 * SB] = fminf(fmaxf(z * inv_aabb + 0.5f, 0.0f), 1.0f);
    out_dir[sid] = 0.5f * dir[r] + 0.5f;
    out_dir[sid + SB] = 0.5f * dir[r + B] + 0.5f;
    out_dir[sid + 2 * SB] = 0.5f * dir[r + 2 * B] + 0.5f;
    out_m_norm[sid] = fminf(fmaxf(((t + 0.5f * dt) - t_min) / t_span, 0.0f), 1.0f);
    out_dt_norm[sid] = dt / t_span;
}

// Next step's ray count: the rays that would fill the sample budget at this
// step's samples-per-ray rate, with 1/8 headroom so the prefix selection
// usually has rays to spare (replaces the reference's host-side estimate,
// trainer.cu:133-140).
__global__ void estimate_rays_kernel(const uint32_t* __restrict__ info, uint32_t max_samples, uint32_t cap,
                                     uint32_t min_rays, uint32_t* __restrict__ n_rays_next) {
    const uint32_t n_rays = info[0], n_samples = info[1];
    uint32_t est = cap;
    if (n_samples > 0 && n_rays > 0) {
        const double per = (double)n_rays / (double)n_samples;
        const double e = per * (double)max_samples * 1.125 + 64.0;
        est = e >= (double)cap ? cap : (uint32_t)e;
    }
    if (est < min_rays) est = min_rays;
    if (est > cap) est = cap;
    *n_rays_next = est;
}

__global__ void accumulate_samples_kernel(const uint32_t* __restrict__ info, unsigned long long* __restrict__ total) {
    total[0] += (unsigned long long)info[1];
}

__global__ void batch_tail_kernel(const uint32_t* __restrict__ info, uint32_t first, uint32_t* __restrict__ tail) {
    const uint32_t n = info[1];
    tail[0] = info[0];
    tail[1] = n > first ? n - first : 0u;
    tail[2] = info[2];
    tail[3] = info[3];
}

}  // namespace tn

using namespace tn;

// total[0] += info[1] on the device (running count of trained samples)
extern "C" int tn_accumulate_samples(const uint32_t* info, unsigned long long* total, hipStream_t stream) {
    TN_CHECK_ARG(info && total);
    accumulate_samples_kernel<<<1, 1, 0, stream>>>(info, total);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_batch_tail(const uint32_t* info, uint32_t first, uint32_t* tail_info, hipStream_t stream) {
    TN_CHECK_ARG(info && tail_info && info != tail_info);
    batch_tail_kernel<<<1, 1, 0, stream>>>(info, first, tail_info);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_march_count(uint32_t n_rays_max, const uint32_t* d_n_rays, uint32_t stride,
                              tn_bbox_t bbox, tn_occupancy_grid_t grid, float cone_angle, float dt_min,
                              float dt_max, uint32_t max_steps, const float* dir, const float* ray_tmax,
                              uint8_t* ray_alive, const float* ori, const float* ray_t, uint32_t* n_steps,
                              hipStream_t stream) {
    TN_CHECK_ARG(dir && ray_tmax && ray_alive && ori && ray_t && n_steps && grid.bitfield);
    TN_CHECK_ARG(n_rays_max <= stride && grid.n_levels >= 1 && grid.n_levels <= 8);
    TN_CHECK_ARG(grid.resolution >= 2 && grid.resolution <= 1024 && dt_min > 0.0f && dt_max >= dt_min);
    if (n_rays_max == 0) return TN_OK;
    march_count_kernel<<<march_blocks(n_rays_max), 64 * kMarchWaves, 0, stream>>>(
        n_rays_max, d_n_rays, stride, bbox, grid, cone_angle, dt_min, dt_max, max_steps, dir, ray_tmax,
        ray_alive, ori, ray_t, n_steps);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_scan_workspace_bytes(uint32_t n) {
    const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
    return (size_t)(tiles == 0 ? 1 : tiles) * sizeof(uint32_t);
}

extern "C" int tn_scan_select(int mode, uint32_t n_rays, const uint32_t* d_n_rays, const uint32_t* n_steps,
                              uint32_t max_samples, uint32_t* ray_offset, uint32_t* info, void* workspace,
                              size_t workspace_bytes, hipStream_t stream) {
    TN_CHECK_ARG(mode == 0 || mode == 1);
    TN_CHECK_ARG(n_steps && ray_offset && info && workspace && n_rays > 0 && max_samples > 0);
    if (workspace_bytes < tn_scan_workspace_bytes(n_rays)) return TN_ERR_OUT_OF_MEMORY;
    const uint32_t tiles = (n_rays + kScanTile - 1) / kScanTile;
    uint32_t* tile_sums = static_cast<uint32_t*>(workspace);
    scan_tile_sums_kernel<<<tiles, kScanThreads, 0, stream>>>(n_rays, d_n_rays, n_steps, tile_sums);
    scan_tile_prefix_kernel<<<1, kScanThreads, 0, stream>>>(tiles, tile_sums);
    scan_write_select_kernel<<<tiles, kScanThreads, 0, stream>>>(n_rays, d_n_rays, n_steps, tile_sums, max_samples,
                                                                 mode, ray_offset, info);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_march_generate(uint32_t n_rays_max, const uint32_t* info, uint32_t stride,
                                 uint32_t sample_stride, tn_bbox_t bbox,
                                 float inv_aabb, tn_occupancy_grid_t grid, float dt_min, float dt_max,
                                 float cone_angle, const uint32_t* ray_img_id, const float* ori,
                                 const float* dir, const float* ray_t, const uint32_t* ray_offset,
                                 const uint8_t* ray_alive, const uint32_t* n_ray_steps, uint32_t* out_img_id,
                                 float* out_t, float* out_pos, float* out_dir, float* out_dt,
                                 float* out_m_norm, float* out_dt_norm, hipStream_t stream) {
    TN_CHECK_ARG(ray_img_id && ori && dir && ray_t && ray_offset && ray_alive && n_ray_steps && out_img_id &&
                 out_t && out_pos && out_dir && out_dt && out_m_norm && out_dt_norm && grid.bitfield);
    TN_CHECK_ARG(n_rays_max <= stride && grid.n_levels >= 1 && grid.n_levels <= 8 && dt_min > 0.0f);
    TN_CHECK_ARG(info != nullptr && sample_stride > 0);
    if (n_rays_max == 0) return TN_OK;
    march_generate_kernel<<<march_blocks(n_rays_max), 64 * kMarchWaves, 0, stream>>>(
        n_rays_max, info, stride, bbox, grid, dt_min, dt_max, cone_angle, ori, dir, ray_t, ray_offset, ray_alive,
        n_ray_steps, out_img_id, out_t);
    march_expand_kernel<<<blocks_for(sample_stride, 256), 256, 0, stream>>>(
        n_rays_max, sample_stride, info, stride, sample_stride, inv_aabb, dt_min, dt_max, cone_angle, ray_img_id, ori,
        dir, ray_t, ray_offset, n_ray_steps, out_img_id, out_t, out_pos, out_dir, out_dt, out_m_norm, out_dt_norm);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_estimate_rays(const uint32_t* info, uint32_t max_samples, uint32_t cap, uint32_t min_rays,
                                uint32_t* n_rays_next, hipStream_t stream) {
    TN_CHECK_ARG(info && n_rays_next && cap > 0);
    estimate_rays_kernel<<<1, 1, 0, stream>>>(info, max_samples, cap, min_rays, n_rays_next);
    TN_CHECK_LAUNCH();
    return TN_OK;
}
