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
    uint32_t gx, gy, gz;  // the cell (its lowest corner)
};

TN_DEVICE void hash_corners(const tn_hashgrid_t& g, int l, float px, float py, float pz, HashCorners& hc) {
    const float s = g.scale[l];
    const float x = px * s + 0.5f, y = py * s + 0.5f, z = pz * s + 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y), fz0 = floorf(z);
    const float fx = x - fx0, fy = y - fy0, fz = z - fz0;
    const uint32_t gx = (uint32_t)(int32_t)fx0, gy = (uint32_t)(int32_t)fy0, gz = (uint32_t)(int32_t)fz0;
    hc.gx = gx;
    hc.gy = gy;
    hc.gz = gz;
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

// Buffer loads of one entry or of the aligned pair of entries e & ~1, e | 1.
// A byte offset at or past the resource's size reads zeros without a memory
// request (the buffer range check), so a lane that needs no load skips it
// without a branch.
template <typename TT>
struct Gather;
template <>
struct Gather<_Float16> {
    using V = half2v;
    typedef uint32_t W __attribute__((ext_vector_type(2)));  // a pair of entries
    TN_DEVICE static W pair(__amdgpu_buffer_rsrc_t r, uint32_t e) {
        return __builtin_amdgcn_raw_buffer_load_b64(r, (int)((e >> 1) * 8u), 0, 0);
    }
    // (the component is copied out first: __builtin_bit_cast of a vector
    // component read the vector's first element in this compiler)
    TN_DEVICE static V even(const W& w) {
        const uint32_t x = w.x;
        return __builtin_bit_cast(V, x);
    }
    TN_DEVICE static V odd(const W& w) {
        const uint32_t y = w.y;
        return __builtin_bit_cast(V, y);
    }
    TN_DEVICE static V one(__amdgpu_buffer_rsrc_t r, uint32_t e, bool need) {
        return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, need ? (int)(e * 4u) : -1, 0, 0));
    }
};
template <>
struct Gather<float> {
    using V = float2;
    typedef uint32_t W __attribute__((ext_vector_type(4)));
    TN_DEVICE static W pair(__amdgpu_buffer_rsrc_t r, uint32_t e) {
        return __builtin_amdgcn_raw_buffer_load_b128(r, (int)((e >> 1) * 16u), 0, 0);
    }
    TN_DEVICE static V even(const W& w) { return make_float2(__uint_as_float(w.x), __uint_as_float(w.y)); }
    TN_DEVICE static V odd(const W& w) { return make_float2(__uint_as_float(w.z), __uint_as_float(w.w)); }
    TN_DEVICE static V one(__amdgpu_buffer_rsrc_t r, uint32_t e, bool need) {
        typedef uint32_t u2v __attribute__((ext_vector_type(2)));
        const u2v w = __builtin_amdgcn_raw_buffer_load_b64(r, need ? (int)(e * 8u) : -1, 0, 0);
        return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
    }
};

// The eight corner entries of a sample-level, read with fewer requests: the
// corners (0, y, z) and (1, y, z) are adjacent entries when the lower one is
// even -- a dense level's x + 1, or a hashed level's even x, whose x + 1 only
// sets the hash's lowest bit -- so one load of the aligned entry pair
// holding the lower corner also yields the upper one; the upper corner is
// loaded on its own only where it lies elsewhere. Same values, and the
// trilinear sums keep their order. `table` spans `bytes` (< 2^32). The
// pairs pass through an empty asm statement once every load is issued:
// without it the compiler narrows each pair load to the half it sees used
// and loads the other corner separately again (eight loads).
template <typename TT>
TN_DEVICE void gather_corners(const TT* __restrict__ table, uint32_t bytes, const HashCorners& hc,
                              typename Pair<TT>::V (&v)[8]) {
    using V = typename Pair<TT>::V;
    using W = typename Gather<TT>::W;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(const_cast<TT*>(table), 0, (int)bytes, 0x00020000);
    W w[4];
    V up[4];
    bool own[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = Gather<TT>::pair(r, hc.idx[2 * q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        own[q] = hc.idx[2 * q + 1] != (hc.idx[2 * q] ^ 1u);
        up[q] = Gather<TT>::one(r, hc.idx[2 * q + 1], own[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        asm volatile("" : "+v"(w[q]));
        const V even = Gather<TT>::even(w[q]), odd = Gather<TT>::odd(w[q]);
        const bool odd0 = (hc.idx[2 * q] & 1u) != 0u;
        v[2 * q] = odd0 ? odd : even;
        v[2 * q + 1] = own[q] ? up[q] : (odd0 ? even : odd);
    }
}

// Levels [l0, l0 + gridDim.y) gathered from global memory.
template <typename TT, typename TO>
__global__ void __launch_bounds__(256) hashgrid_fwd_kernel(uint32_t n_max, const uint32_t* __restrict__ info,
                                                           uint32_t B, uint32_t OB, tn_hashgrid_t g,
                                                           const float* __restrict__ pos, const TT* __restrict__ table,
                                                           TO* __restrict__ out, int l0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int l = l0 + (int)blockIdx.y;
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    if (i >= n) return;
    HashCorners hc;
    hash_corners(g, l, pos[i], pos[i + B], pos[i + 2 * (uint64_t)B], hc);
    // the table's entries (every corner index lies below: levels wrap their indices)
    const uint32_t n_entries = g.level_offset[g.n_levels - 1] + g.level_size[g.n_levels - 1];
    typename Pair<TT>::V v[8];
    gather_corners<TT>(table, n_entries * (uint32_t)(2 * sizeof(TT)), hc, v);
    float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        f0 += hc.w[c] * Pair<TT>::a(v[c]);
        f1 += hc.w[c] * Pair<TT>::b(v[c]);
    }
    out[(uint64_t)(2 * l) * OB + i] = (TO)f0;
    out[(uint64_t)(2 * l + 1) * OB + i] = (TO)f1;
}

// The coarse dense levels [0, Ls) from a copy of their tables in LDS. Their
// entries are the first `stage_bytes` of the table (levels are stored in
// order), small enough to copy into every workgroup: a wave's eight corner
// reads per level then go to LDS instead of each lane requesting its own
// cache line. A workgroup per CU copies the tables once and strides over the
// samples, reading each position once for all Ls levels. The corners, weights
// and the order of the trilinear sums are those of hashgrid_fwd_kernel, so
// the features are the same bit for bit.
constexpr uint32_t kHfLdsThreads = 1024;
constexpr uint32_t kHfLdsBytes = 64 * 1024;
template <typename TT, typename TO>
__global__ void __launch_bounds__(kHfLdsThreads) hashgrid_fwd_lds_kernel(
    uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t OB, tn_hashgrid_t g,
    const float* __restrict__ pos, const TT* __restrict__ table, TO* __restrict__ out, int Ls, uint32_t stage_bytes) {
    extern __shared__ uint4 stage[];
    const uint4* src = reinterpret_cast<const uint4*>(table);
    for (uint32_t k = threadIdx.x; k < stage_bytes / 16u; k += kHfLdsThreads) stage[k] = src[k];
    __syncthreads();
    const auto* st = reinterpret_cast<const typename Pair<TT>::V*>(stage);
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    for (uint32_t i = blockIdx.x * kHfLdsThreads + threadIdx.x; i < n; i += gridDim.x * kHfLdsThreads) {
        const float px = pos[i], py = pos[i + B], pz = pos[i + 2 * (uint64_t)B];
        for (int l = 0; l < Ls; ++l) {
            HashCorners hc;
            hash_corners(g, l, px, py, pz, hc);
            float f0 = 0.0f, f1 = 0.0f;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const typename Pair<TT>::V v = st[hc.idx[c]];
                f0 += hc.w[c] * Pair<TT>::a(v);
                f1 += hc.w[c] * Pair<TT>::b(v);
            }
            out[(uint64_t)(2 * l) * OB + i] = (TO)f0;
            out[(uint64_t)(2 * l + 1) * OB + i] = (TO)f1;
        }
    }
}

// Levels staged in LDS: the longest prefix of dense levels, at most `want`,
// whose entries (entry_bytes each) fit kHfLdsBytes; 0 when the table is not
// 16-byte aligned. *bytes gets the staged size (a multiple of 16: level
// sizes are multiples of 8 entries of at least 4 bytes).
inline int hf_lds_levels(const tn_hashgrid_t& g, uint32_t entry_bytes, const void* table, int want,
                         uint32_t* bytes) {
    *bytes = 0;
    if (reinterpret_cast<uintptr_t>(table) & 15u) return 0;
    int Ls = 0;
    while (Ls < want && Ls < g.n_levels - 1 && g.level_dense[Ls] &&
           (uint64_t)(g.level_offset[Ls] + g.level_size[Ls]) * entry_bytes <= kHfLdsBytes) {
        *bytes = (g.level_offset[Ls] + g.level_size[Ls]) * entry_bytes;
        ++Ls;
    }
    return Ls;
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

// ---------------------------------------------------------------------------
// Binned backward. Global float atomics on this GPU are performed on the
// memory side (none of them hit in L2), and their rate, not bandwidth, bounds
// the scatter-add above. The binned variant turns the scatter into a counting
// sort: the table is cut into bins of 4096 entries; (1) every block counts
// the records it will emit per bin of its level, (2) one block per level
// turns the counts into per-(block, bin) offsets and bin totals, and one
// block scans the totals into bin starts, (3) every block recomputes its
// records, writes them to their bins and notes its largest |value| per bin,
// (4) one block per bin (per chunk of at most kHbChunk records of a large bin)
// sums its records in LDS and adds the sums to the table, (5) the partial
// sums of large bins are added up in a fixed order, so the gradient is the
// same bit for bit from run to run. A record is one sample's contribution to
// both features of one corner entry, after runs of consecutive samples in one
// cell have been merged; samples whose feature gradients are both zero emit
// none. Sums in LDS are 64-bit fixed point (integer LDS atomics run several
// times faster than float ones here): with every |v| < 2^E in the bin and m
// records, the scale 2^S with S = 62 - E - ceil(log2(m + 1)) cannot overflow,
// and the sum of the records is exact up to the rounding of each value to a
// multiple of 2^-S (relative 2^-38 of the bin's largest value or better);
// each record carries its value to 17 significant bits (HbRecord). Blocks of count and
// scatter work on one level each: block b serves level b % L and the
// contiguous 256-sample chunks [j, j + 1) * chunk_blocks, j = b / L, so the
// records of a range of samples can be scattered on their own (the slots of
// every (block, bin) are fixed by the plan) once their gradients exist.
// ---------------------------------------------------------------------------
constexpr int kHbBinLog = 12;
constexpr uint32_t kHbBinEntries = 1u << kHbBinLog;
constexpr uint32_t kHbMaxLevelBins = 512;   // bins per level (level size <= 2^21)
constexpr uint32_t kHbMaxBins = 8192;       // bins in all (16 levels of 2^21 entries)
constexpr uint32_t kHbChunk = 32768;        // records per accumulating block at most
constexpr uint32_t kHbBlocksPerLevel = 256; // count/scatter blocks per level at most

struct HbPlan {
    uint32_t base[TN_HASHGRID_MAX_LEVELS];   // first bin of level l
    uint32_t count[TN_HASHGRID_MAX_LEVELS];  // bins of level l
    uint32_t n_bins;
    uint32_t row;         // per-block stride of the count/offset matrices (max bins of a level)
    uint32_t blocks;      // count/scatter grid (a multiple of the level count)
    uint32_t level_blocks; // count/scatter blocks per level: blocks / L
    uint32_t chunk_blocks; // 256-sample chunks per count/scatter block (block j: chunks [j, j + 1) * chunk_blocks)
    uint32_t acc_blocks;  // accumulate grid: an upper bound on the number of chunks
    uint32_t split_slots; // partial-sum slots for bins summed by several blocks (upper bound)
    uint64_t cap;         // record capacity: n_max * L * 8
};

// A record is one contribution to both features of one entry of a bin, in 8
// bytes: each value rounded to 17 significant bits (relative error at most
// 2^-18, far below the fp16 gradient accumulation of the reference) and the
// bin-relative 12-bit entry index in the freed low mantissa bits, its bits
// 0..5 in the first value's and 6..11 in the second's. (The scatter's write
// and the accumulation's read of the records bound both passes; 12-byte
// records took 1.5 times as long.)
struct HbRecord {
    uint32_t a, b;
};

// v rounded to 17 significant bits (half away from zero; a carry may raise the
// exponent, up to inf), low 6 bits cleared; inf and NaN keep their pattern
TN_DEVICE uint32_t hb_round17(float v) {
    const uint32_t u = __float_as_uint(v);
    if ((u & 0x7F800000u) == 0x7F800000u) return u & ~63u;
    return (u + 32u) & ~63u;
}
TN_DEVICE HbRecord hb_pack(uint32_t e, uint32_t r0, uint32_t r1) { return HbRecord{r0 | (e & 63u), r1 | (e >> 6)}; }
TN_DEVICE uint32_t hb_entry(const HbRecord& r) { return (r.a & 63u) | ((r.b & 63u) << 6); }
TN_DEVICE float hb_v0(const HbRecord& r) { return __uint_as_float(r.a & ~63u); }
TN_DEVICE float hb_v1(const HbRecord& r) { return __uint_as_float(r.b & ~63u); }

// One sample per lane, level l: the eight corner contributions of both
// features, with runs of consecutive lanes on one entry summed into the run's
// last lane (emit). Lanes past the batch and samples whose feature gradients
// are both zero emit nothing.
struct HbCorners {
    uint32_t idx[8];
    float a0[8], a1[8];
    bool emit;
};

// DPP segmented inclusive sum over the wave of lanes with equal key (runs are
// contiguous): rows of 16 first (row_shr 1, 2, 4, 8), then the last lane of
// rows 0 and 2 into rows 1 and 3 (row_bcast:15) and of row 1 into rows 2 and
// 3 (row_bcast:31). A lane adds the shifted value only when the source lane
// has its key; sources outside the row or rows not selected read key 0,
// which no run has. Here for the eight corners' two features under one key.
template <int CTRL, int ROWS>
TN_DEVICE void seg_step16(float (&a0)[8], float (&a1)[8], int key) {
    const bool take = __builtin_amdgcn_update_dpp(0, key, CTRL, ROWS, 0xF, false) == key;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float o0 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a0[c]), CTRL, ROWS, 0xF, false));
        const float o1 = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(a1[c]), CTRL, ROWS, 0xF, false));
        if (take) {
            a0[c] += o0;
            a1[c] += o1;
        }
    }
}

TN_DEVICE void seg_sum16(float (&a0)[8], float (&a1)[8], int key) {
    seg_step16<0x111, 0xF>(a0, a1, key);
    seg_step16<0x112, 0xF>(a0, a1, key);
    seg_step16<0x114, 0xF>(a0, a1, key);
    seg_step16<0x118, 0xF>(a0, a1, key);
    seg_step16<0x142, 0xA>(a0, a1, key);
    seg_step16<0x143, 0xC>(a0, a1, key);
}

// Corner entries and weights of an active sample i at level l, and the mask
// of lanes that start a run: consecutive samples in one cell share all eight
// corner entries, and a run of them (keyed by its ordinal, since a cell may
// recur further on) emits one record per corner from its last lane. Inactive
// lanes carry a cell no sample has, so they stay runs of their own.
TN_DEVICE void hb_cells(const tn_hashgrid_t& g, int l, float x, float y, float z, bool active, HashCorners& c8,
                        uint64_t& heads) {
    const uint32_t lane = __lane_id();
    if (active) {
        hash_corners(g, l, x, y, z, c8);
    } else {
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            c8.idx[c] = 0;
            c8.w[c] = 0.0f;
        }
    }
    const uint32_t cx = active ? c8.gx : 0xFFFFF000u + lane, cy = active ? c8.gy : 0u, cz = active ? c8.gz : 0u;
    const uint32_t px = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)cx, 0x138, 0xF, 0xF, false);  // wave_shr:1
    const uint32_t py = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)cy, 0x138, 0xF, 0xF, false);
    const uint32_t pz = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)cz, 0x138, 0xF, 0xF, false);
    heads = __ballot(lane == 0 || px != cx || py != cy || pz != cz);
}

// Inputs of sample i at level l: its position and both feature gradients
// (zero past the batch). The five loads are independent, so a block loads
// them for its next chunk while it works on the current one.
struct HbIn {
    float x, y, z, g0, g1;
};

template <typename TG>
TN_DEVICE HbIn hb_load(int l, uint32_t i, uint32_t n, uint32_t B, uint32_t GB, const float* __restrict__ pos,
                       const TG* __restrict__ dfeat) {
    HbIn v = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < n) {
        v.x = pos[i];
        v.y = pos[i + B];
        v.z = pos[i + 2 * (uint64_t)B];
        v.g0 = (float)dfeat[(uint64_t)(2 * l) * GB + i];
        v.g1 = (float)dfeat[(uint64_t)(2 * l + 1) * GB + i];
    }
    return v;
}

// every sample of the batch takes part (one with a zero gradient emits
// zero-valued records), so the record counts follow from the positions
// alone and can be formed before the gradients exist
TN_DEVICE bool hb_active(uint32_t i, uint32_t n) { return i < n; }

// position of sample i (zero past the batch), for the count pass
TN_DEVICE HbIn hb_load_pos(uint32_t i, uint32_t n, uint32_t B, const float* __restrict__ pos) {
    HbIn v = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (i < n) {
        v.x = pos[i];
        v.y = pos[i + B];
        v.z = pos[i + 2 * (uint64_t)B];
    }
    return v;
}

// whether this lane ends its run (emits the run's records)
TN_DEVICE bool hb_emits(bool active, uint64_t heads) {
    const uint32_t lane = __lane_id();
    return active && (lane == 63 || ((heads >> (lane + 1)) & 1ull) != 0);
}

TN_DEVICE void hb_corners(const tn_hashgrid_t& g, int l, bool active, const HbIn& v, HbCorners& hc) {
    const uint32_t lane = __lane_id();
    hc.emit = false;
    if (__ballot(active) == 0) return;  // a wave of samples without gradient emits nothing
    HashCorners c8;
    uint64_t hm;
    hb_cells(g, l, v.x, v.y, v.z, active, c8, hm);
    float a0[8], a1[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        a0[c] = c8.w[c] * v.g0;
        a1[c] = c8.w[c] * v.g1;
    }
    if (hm != ~0ull) {
        const uint64_t le = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
        seg_sum16(a0, a1, __builtin_popcountll(hm & le));
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        hc.idx[c] = c8.idx[c];
        hc.a0[c] = a0[c];
        hc.a1[c] = a1[c];
    }
    hc.emit = hb_emits(active, hm);
}

TN_DEVICE uint32_t abs_bits(float a, float b) {
    return max(__float_as_uint(fabsf(a)), __float_as_uint(fabsf(b)));
}

// (1) per-block record counts per bin of the block's level (the run
// structure of the positions only: no gradients are read)
__global__ void __launch_bounds__(256) hashgrid_bwd_count_kernel(uint32_t n_max, const uint32_t* __restrict__ info,
                                                                 uint32_t B, tn_hashgrid_t g, HbPlan p,
                                                                 const float* __restrict__ pos,
                                                                 uint32_t* __restrict__ counts) {
    __shared__ uint32_t hist[kHbMaxLevelBins];
    const uint32_t L = (uint32_t)g.n_levels;
    const int l = (int)(blockIdx.x % L);
    const uint32_t nb = p.count[l];
    for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t c0 = (blockIdx.x / L) * p.chunk_blocks;
    const uint32_t c1 = min(c0 + p.chunk_blocks, (n + 255) / 256);
    HbIn next = hb_load_pos(c0 * 256 + threadIdx.x, n, B, pos);
    for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t i = c * 256 + threadIdx.x;
        const HbIn v = next;
        next = hb_load_pos((c + 1) * 256 + threadIdx.x, n, B, pos);
        const bool active = hb_active(i, n);
        if (__ballot(active) == 0) continue;
        HashCorners c8;
        uint64_t hm;
        hb_cells(g, l, v.x, v.y, v.z, active, c8, hm);
        if (hb_emits(active, hm)) {
#pragma unroll
            for (int k = 0; k < 8; ++k) atomicAdd(&hist[(c8.idx[k] - g.level_offset[l]) >> kHbBinLog], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) counts[(uint64_t)blockIdx.x * p.row + k] = hist[k];
}

// (2a) one block per (level, slice of 64 bins): bin totals and each block's
// record offset within its bins (blocks in order). Thread (q, c) sums its
// rows (at most 16) of column c (a bin) of the level's count matrix, the 16
// row groups of a column are scanned through LDS, and each thread writes the
// offsets of its rows.
constexpr uint32_t kHbReduceRows = 16;
static_assert(kHbBlocksPerLevel <= 16 * kHbReduceRows, "rows per thread");
__global__ void __launch_bounds__(1024) hashgrid_bwd_reduce_kernel(tn_hashgrid_t g, HbPlan p,
                                                                   const uint32_t* __restrict__ counts,
                                                                   uint32_t* __restrict__ offsets,
                                                                   uint32_t* __restrict__ totals) {
    __shared__ uint32_t part[16][64];
    const uint32_t L = (uint32_t)g.n_levels, l = blockIdx.x % L, k0 = (blockIdx.x / L) * 64, nb = p.count[l];
    if (k0 >= nb) return;
    const uint32_t c = threadIdx.x & 63u, q = threadIdx.x >> 6;
    const uint32_t rows = (p.level_blocks + 15) / 16, j0 = q * rows, j1 = min(p.level_blocks, j0 + rows);
    const bool col = k0 + c < nb;
    uint32_t v[kHbReduceRows], sum = 0;
#pragma unroll
    for (uint32_t r = 0; r < kHbReduceRows; ++r) {
        const uint32_t j = j0 + r;
        v[r] = col && j < j1 ? counts[(uint64_t)(l + L * j) * p.row + k0 + c] : 0u;
        sum += v[r];
    }
    part[q][c] = sum;
    __syncthreads();
    uint32_t run = 0;
    for (uint32_t t = 0; t < q; ++t) run += part[t][c];
    if (!col) return;
#pragma unroll
    for (uint32_t r = 0; r < kHbReduceRows; ++r) {
        const uint32_t j = j0 + r;
        if (j < j1) offsets[(uint64_t)(l + L * j) * p.row + k0 + c] = run;
        run += v[r];
    }
    if (q == 15) totals[p.base[l] + k0 + c] = run;
}

// (2b) one block of 1024: exclusive scans of the bin totals (record starts),
// of the chunks per bin (chunk starts) and, for bins summed by several
// blocks, of their chunks (partial-sum slots) and of the bins themselves (the
// list the finishing kernel walks; its length goes to split_list[n_bins]).
// Thread j scans bins [8j, 8j + 8) serially, then the block scans the sums.
constexpr uint32_t kHbScanPer = kHbMaxBins / 1024;
__global__ void __launch_bounds__(1024) hashgrid_bwd_scan_kernel(HbPlan p, const uint32_t* __restrict__ totals,
                                                                 uint32_t* __restrict__ start,
                                                                 uint32_t* __restrict__ chunk_start,
                                                                 uint32_t* __restrict__ slot_base,
                                                                 uint32_t* __restrict__ split_list) {
    __shared__ uint32_t wsum[4][16];
    const uint32_t k0 = kHbScanPer * threadIdx.x;
    const uint32_t lane = __lane_id(), wid = threadIdx.x >> 6;
    uint32_t t[kHbScanPer], c[kHbScanPer], sum[4] = {0, 0, 0, 0};
#pragma unroll
    for (uint32_t j = 0; j < kHbScanPer; ++j) {
        const uint32_t k = k0 + j;
        t[j] = k < p.n_bins ? totals[k] : 0u;
        c[j] = k < p.n_bins ? max(1u, (t[j] + kHbChunk - 1) / kHbChunk) : 0u;
        sum[0] += t[j];
        sum[1] += c[j];
        sum[2] += c[j] > 1 ? c[j] : 0u;
        sum[3] += c[j] > 1 ? 1u : 0u;
    }
    // inclusive scans within each wave, then across the 16 wave totals
    uint32_t inc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        inc[q] = sum[q];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc[q], d, 64);
            if (lane >= (uint32_t)d) inc[q] += o;
        }
        if (lane == 63) wsum[q][wid] = inc[q];
    }
    __syncthreads();
    uint32_t ex[4], all[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t pre = 0, tot = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            const uint32_t v = wsum[q][w];
            pre += w < wid ? v : 0u;
            tot += v;
        }
        ex[q] = pre + inc[q] - sum[q];
        all[q] = tot;
    }
#pragma unroll
    for (uint32_t j = 0; j < kHbScanPer; ++j) {
        const uint32_t k = k0 + j;
        if (k < p.n_bins) {
            start[k] = ex[0];
            chunk_start[k] = ex[1];
            slot_base[k] = ex[2];
            if (c[j] > 1) split_list[ex[3]] = k;
        }
        ex[0] += t[j];
        ex[1] += c[j];
        ex[2] += c[j] > 1 ? c[j] : 0u;
        ex[3] += c[j] > 1 ? 1u : 0u;
    }
    if (threadIdx.x == 0) {
        start[p.n_bins] = all[0];
        chunk_start[p.n_bins] = all[1];
        split_list[p.n_bins] = all[3];
    }
}

// (3) records to their bins and the block's largest |value| per bin into
// column j (the block's index within its level) of that bin's row of binmax
// [n_bins][blocks / L]. A chunk's records (at most 8 per sample) are first
// sorted by bin in LDS -- per-bin ranks from LDS counters, an exclusive scan
// of the counts, then one store each to its place -- and then leave in that
// order, so consecutive lanes write consecutive slots of one bin: the global
// stores go out as runs of whole lines instead of one 8-byte request per
// record, whose rate bounded the unsorted version. The order of records
// within a bin depends on the order of the LDS atomics; the accumulation sums
// them exactly (fixed point), so the result does not.
constexpr uint32_t kHbScatterThreads = 256;
constexpr uint32_t kHbChunkRecords = 8 * kHbScatterThreads;
static_assert(kHbMaxLevelBins <= 2 * kHbScatterThreads, "two bins per thread in the chunk scan");
template <typename TG>
__global__ void __launch_bounds__(kHbScatterThreads) hashgrid_bwd_scatter_kernel(
    uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t GB, tn_hashgrid_t g, HbPlan p,
    const float* __restrict__ pos, const TG* __restrict__ dfeat, const uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ start, uint32_t* __restrict__ binmax, HbRecord* __restrict__ rec, uint32_t l0,
    uint32_t Lg, uint32_t j0) {
    // cur: the bin's next global slot; cnt: the chunk's records per bin;
    // lstart: the chunk's first LDS place of the bin; gdel: global slot minus
    // LDS place for the bin's records of the chunk
    __shared__ uint32_t cur[kHbMaxLevelBins], bmax[kHbMaxLevelBins], cnt[kHbMaxLevelBins];
    __shared__ uint32_t lstart[kHbMaxLevelBins], gdel[kHbMaxLevelBins];
    __shared__ uint2 srec[kHbChunkRecords];
    __shared__ uint32_t sslot[kHbChunkRecords];
    __shared__ uint32_t wtot[kHbScatterThreads / 64];
    const uint32_t L = (uint32_t)g.n_levels;
    // this launch covers levels [l0, l0 + Lg) and count blocks j >= j0;
    // block (l, j) is block l + L j of the count pass
    const int l = (int)(l0 + blockIdx.x % Lg);
    const uint32_t j = j0 + blockIdx.x / Lg, b = (uint32_t)l + L * j;
    const uint32_t nb = p.count[l];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), wid = tid >> 6;
    for (uint32_t k = tid; k < nb; k += kHbScatterThreads) {
        cur[k] = start[p.base[l] + k] + offsets[(uint64_t)b * p.row + k];
        bmax[k] = 0u;
        cnt[k] = 0u;
    }
    const uint32_t off = g.level_offset[l];
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t c0 = j * p.chunk_blocks, c1 = min(c0 + p.chunk_blocks, (n + 255) / 256);
    HbIn next = hb_load<TG>(l, c0 * 256 + tid, n, B, GB, pos, dfeat);
    for (uint32_t c = c0; c < c1; ++c) {
        __syncthreads();  // cnt cleared, srec and wtot free
        const uint32_t i = c * 256 + tid;
        const HbIn v = next;
        next = hb_load<TG>(l, (c + 1) * 256 + tid, n, B, GB, pos, dfeat);
        HbCorners hc;
        hb_corners(g, l, hb_active(i, n), v, hc);
        uint32_t bin[8], rank[8], r0[8], r1[8];
        if (hc.emit) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                bin[k] = (hc.idx[k] - off) >> kHbBinLog;
                rank[k] = atomicAdd(&cnt[bin[k]], 1u);
                r0[k] = hb_round17(hc.a0[k]);
                r1[k] = hb_round17(hc.a1[k]);
                const uint32_t m = max(r0[k] & 0x7FFFFFFFu, r1[k] & 0x7FFFFFFFu);  // of the rounded values
                if (m != 0u) atomicMax(&bmax[bin[k]], m);
            }
        }
        __syncthreads();
        // exclusive scan of the chunk's counts: thread t takes bins 2t, 2t + 1
        const uint32_t k0 = 2 * tid, k1 = k0 + 1;
        const uint32_t c0 = k0 < nb ? cnt[k0] : 0u, c1 = k1 < nb ? cnt[k1] : 0u;
        uint32_t inc = c0 + c1;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d) inc += o;
        }
        if (lane == 63) wtot[wid] = inc;
        __syncthreads();
        uint32_t pre = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < kHbScatterThreads / 64; ++w) {
            const uint32_t t = wtot[w];
            pre += w < wid ? t : 0u;
            total += t;
        }
        const uint32_t e0 = pre + inc - c0 - c1;
        if (k0 < nb) {
            lstart[k0] = e0;
            gdel[k0] = cur[k0] - e0;
            cur[k0] += c0;
            cnt[k0] = 0u;
        }
        if (k1 < nb) {
            lstart[k1] = e0 + c0;
            gdel[k1] = cur[k1] - (e0 + c0);
            cur[k1] += c1;
            cnt[k1] = 0u;
        }
        __syncthreads();
        if (hc.emit) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint32_t q = lstart[bin[k]] + rank[k];
                const HbRecord r = hb_pack((hc.idx[k] - off) & (kHbBinEntries - 1u), r0[k], r1[k]);
                srec[q] = make_uint2(r.a, r.b);
                sslot[q] = q + gdel[bin[k]];
            }
        }
        __syncthreads();
        for (uint32_t q = tid; q < total; q += kHbScatterThreads) {
            const uint32_t slot = sslot[q];
            const uint2 r = srec[q];
            if (slot < p.cap) reinterpret_cast<uint2*>(rec)[slot] = r;  // rec is 8-byte aligned
        }
    }
    __syncthreads();
    for (uint32_t k = tid; k < nb; k += kHbScatterThreads)
        binmax[(uint64_t)(p.base[l] + k) * p.level_blocks + j] = bmax[k];
}

// the largest |value| bits of bin k: the maximum over its row of binmax
TN_DEVICE uint32_t hb_bin_max(const HbPlan& p, const uint32_t* __restrict__ binmax, uint32_t k, uint32_t* smax) {
    if (threadIdx.x == 0) *smax = 0u;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < p.level_blocks; j += blockDim.x) {
        const uint32_t v = binmax[(uint64_t)k * p.level_blocks + j];
        if (v != 0u) atomicMax(smax, v);
    }
    __syncthreads();
    return *smax;
}

TN_DEVICE long long hb_fixed(float v, int S) { return (long long)rintf(ldexpf(v, S)); }

// the complete sums (v0, v1) of table entry `entry` added to dtable
TN_DEVICE void hb_finish_entry(float* __restrict__ dtable, uint32_t entry, float v0, float v1) {
    float2* d = reinterpret_cast<float2*>(dtable) + entry;
    float2 o = *d;
    o.x += v0;
    o.y += v1;
    *d = o;
}

// NGP Adam on the hash table inside the sum passes (tn_adam_t on the device;
// w null: the passes add the gradient to dtable instead)
struct HbAdam {
    float inv_scale0;  // 1 / loss_scale
    const uint32_t* info;
    float lr0;
    const uint32_t* d_step;
    float lr_decay, beta1, beta2, eps;
    float* w;
    _Float16* w_lp;  // null: no fp16 copy
    float* m;
    float* v;
    uint32_t* steps;
};

typedef float hb_f2 __attribute__((ext_vector_type(2)));
typedef uint32_t hb_u2 __attribute__((ext_vector_type(2)));

// Adam on table entry `entry` (its parameters 2 entry and 2 entry + 1) given
// the raw (loss-scaled) gradient sums (s0, s1), by the rule of tn_adam_step:
// g = s * inv, and a zero g leaves its parameter, moments and step untouched.
// Lanes act in groups of 16 consecutive entries (a 128-byte line of each fp32
// array): a group with any gradient reads and writes all its entries back,
// the untouched ones unchanged bit for bit, as adam_kernel does with its
// lines. Every lane of the wave calls it (valid: the lane has an entry).
// Master weights, moments and steps are streamed non-temporally.
TN_DEVICE void hb_adam_entry(const HbAdam& a, float inv, float lr, uint32_t entry, float s0, float s1, bool valid) {
    const float g[2] = {valid ? s0 * inv : 0.0f, valid ? s1 * inv : 0.0f};
    const uint64_t touched = __ballot(g[0] != 0.0f || g[1] != 0.0f);
    const bool group = ((touched >> (__lane_id() & 48u)) & 0xFFFFull) != 0ull;
    if (!valid || !group) return;
    hb_f2* wp = reinterpret_cast<hb_f2*>(a.w) + entry;
    hb_f2* mp = reinterpret_cast<hb_f2*>(a.m) + entry;
    hb_f2* vp = reinterpret_cast<hb_f2*>(a.v) + entry;
    hb_u2* sp = reinterpret_cast<hb_u2*>(a.steps) + entry;
    const hb_f2 w2 = __builtin_nontemporal_load(wp), m2 = __builtin_nontemporal_load(mp);
    const hb_f2 v2 = __builtin_nontemporal_load(vp);
    const hb_u2 s2 = __builtin_nontemporal_load(sp);
    float w[2] = {w2.x, w2.y}, m[2] = {m2.x, m2.y}, v[2] = {v2.x, v2.y};
    uint32_t st[2] = {s2.x, s2.y};
#pragma unroll
    for (int f = 0; f < 2; ++f) {
        if (g[f] == 0.0f) continue;
        const uint32_t t = st[f] + 1;
        adam_apply(g[f], adam_bias(t, a.beta1, a.beta2), lr, a.beta1, a.beta2, a.eps, w[f], m[f], v[f]);
        st[f] = t;
    }
    __builtin_nontemporal_store(hb_f2{w[0], w[1]}, wp);
    __builtin_nontemporal_store(hb_f2{m[0], m[1]}, mp);
    __builtin_nontemporal_store(hb_f2{v[0], v[1]}, vp);
    __builtin_nontemporal_store(hb_u2{st[0], st[1]}, sp);
    if (a.w_lp) {
        half2v h;
        h.x = (_Float16)w[0];
        h.y = (_Float16)w[1];
        reinterpret_cast<half2v*>(a.w_lp)[entry] = h;
    }
}

typedef float hb_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t hb_u4 __attribute__((ext_vector_type(4)));
typedef _Float16 hb_h4 __attribute__((ext_vector_type(4)));

// The same rule on P pairs of entries per lane, pair u being entries
// ent[u], ent[u] + 1 (ent[u] even, so its four parameters are one 16-byte
// access of each array) with raw sums s[u][0..3]; groups of 8 lanes (16
// entries, a 128-byte line of each fp32 array) act together. Every pair's
// loads are issued before any update, so a lane keeps P lines in flight.
template <int P>
TN_DEVICE void hb_adam_pairs(const HbAdam& a, float inv, float lr, const uint32_t (&ent)[P], const float (&s)[P][4],
                             const bool (&valid)[P]) {
    float g[P][4];
    bool act[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
        bool any = false;
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            g[u][f] = valid[u] ? s[u][f] * inv : 0.0f;
            any |= g[u][f] != 0.0f;
        }
        const uint64_t touched = __ballot(any);
        act[u] = valid[u] && ((touched >> (__lane_id() & 56u)) & 0xFFull) != 0ull;
    }
    hb_f4 w[P], m[P], v[P];
    hb_u4 st[P];
#pragma unroll
    for (int u = 0; u < P; ++u) {
        if (!act[u]) continue;
        const uint32_t q = ent[u] >> 1;
        w[u] = __builtin_nontemporal_load(reinterpret_cast<const hb_f4*>(a.w) + q);
        m[u] = __builtin_nontemporal_load(reinterpret_cast<const hb_f4*>(a.m) + q);
        v[u] = __builtin_nontemporal_load(reinterpret_cast<const hb_f4*>(a.v) + q);
        st[u] = __builtin_nontemporal_load(reinterpret_cast<const hb_u4*>(a.steps) + q);
    }
#pragma unroll
    for (int u = 0; u < P; ++u) {
        if (!act[u]) continue;
        const uint32_t q = ent[u] >> 1;
        float wv[4] = {w[u].x, w[u].y, w[u].z, w[u].w}, mv[4] = {m[u].x, m[u].y, m[u].z, m[u].w};
        float vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
        uint32_t sv[4] = {st[u].x, st[u].y, st[u].z, st[u].w};
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            if (g[u][f] == 0.0f) continue;
            const uint32_t t = sv[f] + 1;
            adam_apply(g[u][f], adam_bias(t, a.beta1, a.beta2), lr, a.beta1, a.beta2, a.eps, wv[f], mv[f], vv[f]);
            sv[f] = t;
        }
        __builtin_nontemporal_store(hb_f4{wv[0], wv[1], wv[2], wv[3]}, reinterpret_cast<hb_f4*>(a.w) + q);
        __builtin_nontemporal_store(hb_f4{mv[0], mv[1], mv[2], mv[3]}, reinterpret_cast<hb_f4*>(a.m) + q);
        __builtin_nontemporal_store(hb_f4{vv[0], vv[1], vv[2], vv[3]}, reinterpret_cast<hb_f4*>(a.v) + q);
        __builtin_nontemporal_store(hb_u4{sv[0], sv[1], sv[2], sv[3]}, reinterpret_cast<hb_u4*>(a.steps) + q);
        if (a.w_lp)
            reinterpret_cast<hb_h4*>(a.w_lp)[q] =
                hb_h4{(_Float16)wv[0], (_Float16)wv[1], (_Float16)wv[2], (_Float16)wv[3]};
    }
}

TN_DEVICE float hb_adam_inv(const HbAdam& a) { return a.info ? a.inv_scale0 / ray_pow2(a.info[0]) : a.inv_scale0; }

// bin k's level, first entry and entry count
struct HbBin {
    uint32_t l, local0, ne;
};
TN_DEVICE HbBin hb_bin(const tn_hashgrid_t& g, const HbPlan& p, uint32_t k) {
    HbBin b;
    b.l = 0;
    while (b.l + 1 < (uint32_t)g.n_levels && p.base[b.l + 1] <= k) ++b.l;
    b.local0 = (k - p.base[b.l]) << kHbBinLog;
    b.ne = min(kHbBinEntries, g.level_size[b.l] - b.local0);
    return b;
}

// fixed-point scale of bin k: every |v| < 2^E with m records in the bin
TN_DEVICE int hb_scale(uint32_t bits, uint64_t m) {
    const int E = (int)(bits >> 23) - 126;
    return 62 - E - (32 - __builtin_clz((uint32_t)m + 1u));
}

// (4) one block per chunk of a bin: fixed-point sums in LDS. A bin summed by
// one block adds its sums to dtable here; a bin split over several
// blocks leaves each block's exact 64-bit partial sums in its slot for (5),
// so the result does not depend on the order the blocks run in. A bin
// holding a non-finite value sums in float and adds into dtable atomically
// instead, so inf and NaN reach the table.
__global__ void __launch_bounds__(512) hashgrid_bwd_accumulate_kernel(tn_hashgrid_t g, HbPlan p,
                                                                      const uint32_t* __restrict__ start,
                                                                      const uint32_t* __restrict__ chunk_start,
                                                                      const uint32_t* __restrict__ binmax,
                                                                      const uint32_t* __restrict__ slot_base,
                                                                      const HbRecord* __restrict__ rec,
                                                                      unsigned long long* __restrict__ slots,
                                                                      float* __restrict__ dtable, uint32_t k_lo,
                                                                      uint32_t k_hi, HbAdam adam) {
    __shared__ unsigned long long acc[2 * kHbBinEntries];
    __shared__ uint32_t smax;
    // this launch covers the chunks of bins [k_lo, k_hi)
    const uint32_t j = chunk_start[k_lo] + blockIdx.x;
    if (j >= chunk_start[k_hi]) return;
    uint32_t lo = k_lo, hi = k_hi;  // the bin k with chunk_start[k] <= j < chunk_start[k + 1]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (chunk_start[mid] <= j) lo = mid;
        else hi = mid;
    }
    const uint32_t k = lo;
    const uint32_t part = j - chunk_start[k], nparts = chunk_start[k + 1] - chunk_start[k];
    const HbBin bin = hb_bin(g, p, k);
    const uint32_t ne = bin.ne;
    for (uint32_t e = threadIdx.x; e < 2 * ne; e += blockDim.x) acc[e] = 0ull;
    const uint32_t bits = hb_bin_max(p, binmax, k, &smax);  // its barriers also order the clearing
    const uint64_t b0 = start[k], b1 = min((uint64_t)start[k + 1], p.cap);
    const uint64_t r0 = b0 + (uint64_t)part * kHbChunk;
    const uint64_t r1 = min(b1, r0 + kHbChunk);
    const bool finite = bits < 0x7F800000u;
    const int S = hb_scale(bits, b1 > b0 ? b1 - b0 : 0);
    float* accf = reinterpret_cast<float*>(acc);
    constexpr int kRecUnroll = 8;  // records in flight per lane
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += kRecUnroll * blockDim.x) {
        HbRecord q[kRecUnroll];
        bool in[kRecUnroll];
#pragma unroll
        for (int u = 0; u < kRecUnroll; ++u) {
            const uint64_t rr = r + (uint64_t)u * blockDim.x;
            in[u] = rr < r1;
            q[u] = in[u] ? rec[rr] : HbRecord{0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < kRecUnroll; ++u) {
            const uint32_t e = hb_entry(q[u]);
            if (!in[u] || e >= ne) continue;
            if (finite) {
                atomicAdd(&acc[2 * e], (unsigned long long)hb_fixed(hb_v0(q[u]), S));
                atomicAdd(&acc[2 * e + 1], (unsigned long long)hb_fixed(hb_v1(q[u]), S));
            } else {
                atomicAdd(&accf[2 * e], hb_v0(q[u]));
                atomicAdd(&accf[2 * e + 1], hb_v1(q[u]));
            }
        }
    }
    __syncthreads();
    const uint32_t entry0 = g.level_offset[bin.l] + bin.local0;
    // all parts of a split bin use slots, or none (the bound on slots covers
    // every batch, so "none" is only a guard)
    const bool to_slots = nparts > 1 && finite && (uint64_t)slot_base[k] + nparts <= p.split_slots;
    if (to_slots) {
        unsigned long long* dst = slots + (uint64_t)(slot_base[k] + part) * 2 * kHbBinEntries;
        for (uint32_t e = threadIdx.x; e < 2 * ne; e += blockDim.x) dst[e] = acc[e];
        return;
    }
    // with Adam, a bin summed by one block takes its step here; a split bin
    // still summed in float adds into dtable for (5) to take
    const bool adam_here = adam.w && nparts == 1;
    const float inv = adam_here ? hb_adam_inv(adam) : 0.0f;
    const float lr = adam_here ? step_lr(adam.lr0, adam.d_step, adam.lr_decay) : 0.0f;
    // lanes take pairs of entries (e, e + 1), two pairs a round: bins and
    // levels start at multiples of 8 entries and ne is one, so pairs never
    // straddle the bin's end and entry0 + e is even
    constexpr int kPairs = 2;
    for (uint32_t e0 = 0; e0 < ne; e0 += 2 * kPairs * blockDim.x) {  // the same trip count in every lane
        uint32_t ent[kPairs];
        float sum[kPairs][4];
        bool valid[kPairs];
#pragma unroll
        for (int u = 0; u < kPairs; ++u) {
            const uint32_t e = e0 + 2 * (u * blockDim.x + threadIdx.x);
            valid[u] = e < ne;
            ent[u] = entry0 + e;
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                float x = 0.0f;
                if (valid[u] && finite) {
                    const long long q = (long long)acc[2 * e + f];
                    if (q != 0) x = (float)ldexp((double)q, -S);
                } else if (valid[u]) {
                    x = accf[2 * e + f];
                }
                sum[u][f] = x;
            }
        }
        if (adam_here) {
            hb_adam_pairs<kPairs>(adam, inv, lr, ent, sum, valid);
            continue;
        }
#pragma unroll
        for (int u = 0; u < kPairs; ++u) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float v0 = sum[u][2 * h], v1 = sum[u][2 * h + 1];
                if (!valid[u] || (v0 == 0.0f && v1 == 0.0f)) continue;
                const uint32_t entry = ent[u] + h;
                if (nparts == 1) {
                    hb_finish_entry(dtable, entry, v0, v1);
                } else {  // a split bin summed in float: partials added atomically
                    atomicAdd(&dtable[2 * entry], v0);
                    atomicAdd(&dtable[2 * entry + 1], v1);
                }
            }
        }
    }
}

// (5) bins split over several blocks: the sum of their partial sums, in slot
// order (exact in 64 bits), finished like a one-block bin. Work items are
// (split bin, 512 entries of it); blocks stride over them.
constexpr uint32_t kHbFinishThreads = 512;
constexpr uint32_t kHbFinishItems = kHbBinEntries / kHbFinishThreads;  // per bin
constexpr uint32_t kHbFinishBlocks = 256;
__global__ void __launch_bounds__(kHbFinishThreads) hashgrid_bwd_finish_kernel(
    tn_hashgrid_t g, HbPlan p, const uint32_t* __restrict__ start, const uint32_t* __restrict__ chunk_start,
    const uint32_t* __restrict__ binmax, const uint32_t* __restrict__ slot_base, const uint32_t* __restrict__ split_list,
    const unsigned long long* __restrict__ slots, float* __restrict__ dtable, HbAdam adam) {
    __shared__ uint32_t smax;
    const uint32_t items = split_list[p.n_bins] * kHbFinishItems;
    const float inv = adam.w ? hb_adam_inv(adam) : 0.0f;
    const float lr = adam.w ? step_lr(adam.lr0, adam.d_step, adam.lr_decay) : 0.0f;
    for (uint32_t w = blockIdx.x; w < items; w += gridDim.x) {
        const uint32_t k = split_list[w / kHbFinishItems];
        const uint32_t bits = hb_bin_max(p, binmax, k, &smax);
        __syncthreads();  // smax is rewritten by the next item
        const uint32_t nparts = chunk_start[k + 1] - chunk_start[k], s0 = slot_base[k];
        // non-finite, or beyond the slots: (4) added the bin into dtable
        // atomically; with Adam its step is taken from there
        const bool in_slots = bits < 0x7F800000u && (uint64_t)s0 + nparts <= p.split_slots;
        if (!in_slots && !adam.w) continue;
        const HbBin bin = hb_bin(g, p, k);
        const uint32_t e = (w % kHbFinishItems) * kHbFinishThreads + threadIdx.x;
        const bool valid = e < bin.ne;
        const uint32_t entry = g.level_offset[bin.l] + bin.local0 + e;
        float v0 = 0.0f, v1 = 0.0f;
        if (valid && in_slots) {
            const uint64_t b0 = start[k], b1 = min((uint64_t)start[k + 1], p.cap);
            const int S = hb_scale(bits, b1 > b0 ? b1 - b0 : 0);
            const ulonglong2* src = reinterpret_cast<const ulonglong2*>(slots) + (uint64_t)s0 * kHbBinEntries + e;
            long long q0 = 0, q1 = 0;
            uint32_t j = 0;
            for (; j + 4 <= nparts; j += 4) {  // four loads in flight, added in slot order
                ulonglong2 t[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) t[u] = src[(uint64_t)(j + u) * kHbBinEntries];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    q0 += (long long)t[u].x;
                    q1 += (long long)t[u].y;
                }
            }
            for (; j < nparts; ++j) {
                const ulonglong2 t = src[(uint64_t)j * kHbBinEntries];
                q0 += (long long)t.x;
                q1 += (long long)t.y;
            }
            if (q0 != 0) v0 = (float)ldexp((double)q0, -S);
            if (q1 != 0) v1 = (float)ldexp((double)q1, -S);
        } else if (valid) {  // with Adam: the bin's float sums, taken out of dtable
            float2* d = reinterpret_cast<float2*>(dtable) + entry;
            const float2 o = *d;
            v0 = o.x;
            v1 = o.y;
            if (v0 != 0.0f || v1 != 0.0f) *d = make_float2(0.0f, 0.0f);
        }
        if (adam.w)
            hb_adam_entry(adam, inv, lr, entry, v0, v1, valid);
        else if (valid && (v0 != 0.0f || v1 != 0.0f))
            hb_finish_entry(dtable, entry, v0, v1);
    }
}

// Bin plan for a layout and batch capacity with at most `level_blocks_max`
// count/scatter workgroups per level; false when the table does not fit the
// bin limits (the caller then takes the atomic kernel).
inline bool hb_plan(const tn_hashgrid_t& g, uint32_t n_max, uint32_t level_blocks_max, HbPlan& p) {
    p = HbPlan{};
    uint32_t n_bins = 0, row = 0;
    for (int l = 0; l < g.n_levels; ++l) {
        p.base[l] = n_bins;
        p.count[l] = (g.level_size[l] + kHbBinEntries - 1) / kHbBinEntries;
        n_bins += p.count[l];
        row = row > p.count[l] ? row : p.count[l];
    }
    if (row > kHbMaxLevelBins || n_bins > kHbMaxBins || n_bins == 0) return false;
    const uint64_t L = (uint64_t)g.n_levels;
    const uint64_t chunks = ((uint64_t)n_max + 255) / 256;
    const uint64_t per_level = chunks < level_blocks_max ? (chunks > 0 ? chunks : 1) : level_blocks_max;
    p.n_bins = n_bins;
    p.row = row;
    p.blocks = (uint32_t)(per_level * L);
    p.level_blocks = (uint32_t)per_level;
    p.chunk_blocks = (uint32_t)((chunks + per_level - 1) / per_level);
    p.cap = (uint64_t)n_max * L * 8;
    if (p.cap >= (1ull << 32)) return false;  // record offsets are 32-bit
    p.acc_blocks = (uint32_t)((uint64_t)n_bins + (p.cap + kHbChunk - 1) / kHbChunk);
    // a bin with t > kHbChunk records has ceil(t / kHbChunk) <= 2 t / kHbChunk chunks
    p.split_slots = (uint32_t)(2 * ((p.cap + kHbChunk - 1) / kHbChunk) + 2);
    return true;
}

// The workspace in two parts: the plan (what the positions decide: record
// counts and offsets per block and bin, bin starts, chunking) and the work
// area of the gradient passes (per-block bin maxima, the records, partial
// sums of split bins).
struct HbPlanArea {
    uint32_t *counts, *offsets, *totals, *start, *chunk_start, *slot_base, *split_list;
};
struct HbWorkArea {
    uint32_t* binmax;
    unsigned long long* slots;
    HbRecord* rec;
};

inline size_t hb_align(size_t x) { return (x + 255) & ~(size_t)255; }

struct HbCarver {
    unsigned char* b;
    size_t off = 0;
    unsigned char* take(size_t bytes) {
        unsigned char* q = b ? b + off : nullptr;
        off += hb_align(bytes);
        return q;
    }
};

// plan bytes, and the carve-up when `base` is given
inline size_t hb_plan_area(const HbPlan& p, void* base, HbPlanArea* pa) {
    const size_t mat = (size_t)p.blocks * p.row * sizeof(uint32_t);
    const size_t bins1 = ((size_t)p.n_bins + 1) * sizeof(uint32_t);
    HbCarver c{static_cast<unsigned char*>(base)};
    HbPlanArea a;
    a.counts = reinterpret_cast<uint32_t*>(c.take(mat));
    a.offsets = reinterpret_cast<uint32_t*>(c.take(mat));
    a.totals = reinterpret_cast<uint32_t*>(c.take(bins1));
    a.start = reinterpret_cast<uint32_t*>(c.take(bins1));
    a.chunk_start = reinterpret_cast<uint32_t*>(c.take(bins1));
    a.slot_base = reinterpret_cast<uint32_t*>(c.take(bins1));
    a.split_list = reinterpret_cast<uint32_t*>(c.take(bins1 + sizeof(uint32_t)));
    if (pa) *pa = a;
    return c.off;
}

// work-area bytes, and the carve-up when `base` is given
inline size_t hb_work_area(const HbPlan& p, void* base, HbWorkArea* wa) {
    HbCarver c{static_cast<unsigned char*>(base)};
    HbWorkArea a;
    a.binmax = reinterpret_cast<uint32_t*>(c.take((size_t)p.n_bins * p.level_blocks * sizeof(uint32_t)));
    a.slots = reinterpret_cast<unsigned long long*>(
        c.take((size_t)p.split_slots * 2 * kHbBinEntries * sizeof(unsigned long long)));
    a.rec = reinterpret_cast<HbRecord*>(c.take((size_t)p.cap * sizeof(HbRecord)));
    if (wa) *wa = a;
    return c.off;
}

// both parts in one buffer (the plan first)
inline size_t hb_workspace(const HbPlan& p) { return hb_plan_area(p, nullptr, nullptr) + hb_work_area(p, nullptr, nullptr); }

// (1)-(2b): the plan from the positions
inline void hashgrid_backward_plan(uint32_t n_max, const uint32_t* info, uint32_t B, const tn_hashgrid_t& g,
                                   const HbPlan& p, const float* pos, const HbPlanArea& pa, hipStream_t stream) {
    hashgrid_bwd_count_kernel<<<p.blocks, 256, 0, stream>>>(n_max, info, B, g, p, pos, pa.counts);
    hashgrid_bwd_reduce_kernel<<<g.n_levels * ((p.row + 63) / 64), 1024, 0, stream>>>(g, p, pa.counts, pa.offsets,
                                                                                     pa.totals);
    hashgrid_bwd_scan_kernel<<<1, 1024, 0, stream>>>(p, pa.totals, pa.start, pa.chunk_start, pa.slot_base,
                                                     pa.split_list);
}

// (3) for levels [l0, l0 + Lg) and the samples of count blocks [j0, j1)
template <typename TG>
void hb_scatter(uint32_t n_max, const uint32_t* info, uint32_t B, uint32_t GB, const tn_hashgrid_t& g, const HbPlan& p,
                const float* pos, const TG* dfeat, const HbPlanArea& pa, const HbWorkArea& wa, uint32_t l0,
                uint32_t Lg, uint32_t j0, uint32_t j1, hipStream_t stream) {
    if (j1 > j0)
        hashgrid_bwd_scatter_kernel<TG><<<(j1 - j0) * Lg, kHbScatterThreads, 0, stream>>>(
            n_max, info, B, GB, g, p, pos, dfeat, pa.offsets, pa.start, wa.binmax, wa.rec, l0, Lg, j0);
}

// (4) for levels [l0, l0 + Lg), every record of them scattered
inline void hb_accumulate(uint32_t n_max, const tn_hashgrid_t& g, const HbPlan& p, float* dtable,
                          const HbPlanArea& pa, const HbWorkArea& wa, uint32_t l0, uint32_t Lg, const HbAdam& adam,
                          hipStream_t stream) {
    const uint32_t L = (uint32_t)g.n_levels;
    const uint32_t k_lo = p.base[l0], k_hi = l0 + Lg < L ? p.base[l0 + Lg] : p.n_bins;
    // chunks of the group's bins: at most one per bin plus one per kHbChunk records
    const uint64_t cap_g = (uint64_t)n_max * Lg * 8;
    const uint32_t acc_blocks = (uint32_t)((k_hi - k_lo) + (cap_g + kHbChunk - 1) / kHbChunk);
    hashgrid_bwd_accumulate_kernel<<<acc_blocks, 512, 0, stream>>>(g, p, pa.start, pa.chunk_start, wa.binmax,
                                                                   pa.slot_base, wa.rec, wa.slots, dtable, k_lo, k_hi,
                                                                   adam);
}

// (5)
inline void hb_finish(const tn_hashgrid_t& g, const HbPlan& p, float* dtable, const HbPlanArea& pa,
                      const HbWorkArea& wa, const HbAdam& adam, hipStream_t stream) {
    hashgrid_bwd_finish_kernel<<<kHbFinishBlocks, kHbFinishThreads, 0, stream>>>(
        g, p, pa.start, pa.chunk_start, wa.binmax, pa.slot_base, pa.split_list, wa.slots, dtable, adam);
}

// (3)-(5): records, sums and the table update, from the plan and the gradients
template <typename TG>
void hashgrid_backward_planned(uint32_t n_max, const uint32_t* info, uint32_t B, uint32_t GB, const tn_hashgrid_t& g,
                               const HbPlan& p, const float* pos, const TG* dfeat, float* dtable,
                               const HbPlanArea& pa, const HbWorkArea& wa, hipStream_t stream,
                               const HbAdam& adam = HbAdam{}) {
    // scatter and accumulate by groups of levels, so a group's records are
    // read back while they are still in the device's last-level cache
    const uint32_t L = (uint32_t)g.n_levels;
    const uint32_t G = std::min<uint32_t>(L, (uint32_t)tune_value(TN_TUNE_HASH_BWD_LEVEL_GROUPS));
    const uint32_t per = (L + G - 1) / G;
    for (uint32_t l0 = 0; l0 < L; l0 += per) {
        const uint32_t Lg = std::min(per, L - l0);
        hb_scatter<TG>(n_max, info, B, GB, g, p, pos, dfeat, pa, wa, l0, Lg, 0, p.level_blocks, stream);
        hb_accumulate(n_max, g, p, dtable, pa, wa, l0, Lg, adam, stream);
    }
    hb_finish(g, p, dtable, pa, wa, adam, stream);
}

inline bool hb_layout_ok(const tn_hashgrid_t& g) {
    return g.n_levels >= 1 && g.n_levels <= TN_HASHGRID_MAX_LEVELS && g.n_features == 2;
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
    if (!((table_dtype == 0 && feat_dtype == 0) || (table_dtype == 1 && (feat_dtype == 0 || feat_dtype == 1))))
        return TN_ERR_INVALID_ARGUMENT;
    const unsigned tb = (unsigned)tune_value(TN_TUNE_HASH_FWD_BLOCK);  // <= 256, the kernel's bound
    uint32_t stage = 0;
    const int Ls = hf_lds_levels(grid, table_dtype == 1 ? 4u : 8u, table,
                                 tune_value(TN_TUNE_HASH_FWD_LDS_LEVELS), &stage);
    if (Ls > 0) {
        const unsigned lb = std::min<unsigned>(blocks_for(n_max, kHfLdsThreads), compute_units());
        if (table_dtype == 0) {
            if (!allow_lds(hashgrid_fwd_lds_kernel<float, float>, stage)) return TN_ERR_LAUNCH;
            hashgrid_fwd_lds_kernel<float, float><<<lb, kHfLdsThreads, stage, stream>>>(
                n_max, info, stride, out_stride, grid, pos, static_cast<const float*>(table),
                static_cast<float*>(features), Ls, stage);
        } else if (feat_dtype == 1) {
            if (!allow_lds(hashgrid_fwd_lds_kernel<_Float16, _Float16>, stage)) return TN_ERR_LAUNCH;
            hashgrid_fwd_lds_kernel<_Float16, _Float16><<<lb, kHfLdsThreads, stage, stream>>>(
                n_max, info, stride, out_stride, grid, pos, static_cast<const _Float16*>(table),
                static_cast<_Float16*>(features), Ls, stage);
        } else {
            if (!allow_lds(hashgrid_fwd_lds_kernel<_Float16, float>, stage)) return TN_ERR_LAUNCH;
            hashgrid_fwd_lds_kernel<_Float16, float><<<lb, kHfLdsThreads, stage, stream>>>(
                n_max, info, stride, out_stride, grid, pos, static_cast<const _Float16*>(table),
                static_cast<float*>(features), Ls, stage);
        }
    }
    const dim3 blocks(blocks_for(n_max, tb), grid.n_levels - Ls);
    if (table_dtype == 0)
        hashgrid_fwd_kernel<float, float><<<blocks, tb, 0, stream>>>(n_max, info, stride, out_stride, grid, pos,
                                                                      static_cast<const float*>(table),
                                                                      static_cast<float*>(features), Ls);
    else if (feat_dtype == 1)
        hashgrid_fwd_kernel<_Float16, _Float16><<<blocks, tb, 0, stream>>>(
            n_max, info, stride, out_stride, grid, pos, static_cast<const _Float16*>(table),
            static_cast<_Float16*>(features), Ls);
    else
        hashgrid_fwd_kernel<_Float16, float><<<blocks, tb, 0, stream>>>(n_max, info, stride, out_stride, grid, pos,
                                                                         static_cast<const _Float16*>(table),
                                                                         static_cast<float*>(features), Ls);
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

// Workspace of tn_hashgrid_backward_binned for batches of up to n_max
// samples (for any TN_TUNE_HASH_BWD_LEVEL_BLOCKS); 0 when the layout exceeds
// the bin limits.
extern "C" size_t tn_hashgrid_backward_workspace_bytes(uint32_t n_max, tn_hashgrid_t grid) {
    HbPlan p;
    if (!hb_layout_ok(grid) || n_max == 0 || !hb_plan(grid, n_max, kHbBlocksPerLevel, p)) return 0;
    return hb_workspace(p);
}

// The scatter-add of tn_hashgrid_backward through the bin sort (same
// arguments and result up to summation order); a missing or short workspace,
// or a layout beyond the bin limits, takes the atomic kernel.
extern "C" int tn_hashgrid_backward_binned(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
                                           uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos,
                                           const void* dfeat, float* dtable, void* workspace, size_t workspace_bytes,
                                           hipStream_t stream) {
    TN_CHECK_ARG(pos && dfeat && dtable && n_max <= stride && n_max <= dfeat_stride && hb_layout_ok(grid));
    TN_CHECK_ARG(dfeat_dtype == 0 || dfeat_dtype == 1);
    if (n_max == 0) return TN_OK;
    HbPlan p;
    const uint32_t level_blocks = (uint32_t)tune_value(TN_TUNE_HASH_BWD_LEVEL_BLOCKS);
    if (!workspace || !hb_plan(grid, n_max, level_blocks, p) || workspace_bytes < hb_workspace(p))
        return tn_hashgrid_backward(dfeat_dtype, n_max, info, stride, dfeat_stride, grid, pos, dfeat, dtable, stream);
    TN_CHECK_ARG((reinterpret_cast<uintptr_t>(workspace) & 255) == 0 && (reinterpret_cast<uintptr_t>(dtable) & 7) == 0);
    HbPlanArea pa;
    HbWorkArea wa;
    const size_t plan_bytes = hb_plan_area(p, workspace, &pa);
    hb_work_area(p, static_cast<unsigned char*>(workspace) + plan_bytes, &wa);
    hashgrid_backward_plan(n_max, info, stride, grid, p, pos, pa, stream);
    if (dfeat_dtype == 0)
        hashgrid_backward_planned<float>(n_max, info, stride, dfeat_stride, grid, p, pos,
                                         static_cast<const float*>(dfeat), dtable, pa, wa, stream);
    else
        hashgrid_backward_planned<_Float16>(n_max, info, stride, dfeat_stride, grid, p, pos,
                                            static_cast<const _Float16*>(dfeat), dtable, pa, wa, stream);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// tn_hashgrid_backward_binned in two calls, for callers that know the
// positions before the gradients: the plan (record counts and offsets per
// bin, from the positions alone) into `plan`, then the gradient passes into
// `work`. Both calls take the batch's positions and bound, and use 256
// count/scatter workgroups per level (not the tuning knob, so that a plan
// made earlier always fits the passes that use it).
static bool hb_split_plan(const tn_hashgrid_t& g, uint32_t n_max, HbPlan& p) {
    return hb_layout_ok(g) && n_max > 0 && hb_plan(g, n_max, kHbBlocksPerLevel, p);
}

extern "C" size_t tn_hashgrid_backward_plan_bytes(uint32_t n_max, tn_hashgrid_t grid) {
    HbPlan p;
    return hb_split_plan(grid, n_max, p) ? hb_plan_area(p, nullptr, nullptr) : 0;
}

extern "C" size_t tn_hashgrid_backward_work_bytes(uint32_t n_max, tn_hashgrid_t grid) {
    HbPlan p;
    return hb_split_plan(grid, n_max, p) ? hb_work_area(p, nullptr, nullptr) : 0;
}

extern "C" int tn_hashgrid_backward_plan(uint32_t n_max, const uint32_t* info, uint32_t stride, tn_hashgrid_t grid,
                                         const float* pos, void* plan, size_t plan_bytes, hipStream_t stream) {
    TN_CHECK_ARG(pos && plan && n_max <= stride);
    HbPlan p;
    TN_CHECK_ARG(n_max == 0 || hb_split_plan(grid, n_max, p));
    if (n_max == 0) return TN_OK;
    TN_CHECK_ARG(plan_bytes >= hb_plan_area(p, nullptr, nullptr) && (reinterpret_cast<uintptr_t>(plan) & 255) == 0);
    HbPlanArea pa;
    hb_plan_area(p, plan, &pa);
    hashgrid_backward_plan(n_max, info, stride, grid, p, pos, pa, stream);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// tn_adam_t on the device; false when its pointers are missing or misaligned
// (16-byte accesses of four parameters, 8 bytes of the fp16 copy)
static bool hb_adam_from(const tn_adam_t* a, HbAdam& h) {
    auto al = [](const void* q, uintptr_t n) { return q && (reinterpret_cast<uintptr_t>(q) & (n - 1)) == 0; };
    if (!a || !(a->loss_scale > 0.0f) || (a->lp_dtype != 0 && a->lp_dtype != 1)) return false;
    if (!al(a->weights, 16) || !al(a->first_moments, 16) || !al(a->second_moments, 16) || !al(a->param_steps, 16) ||
        (a->lp_dtype == 1 && !al(a->weights_lp, 8)))
        return false;
    h = HbAdam{1.0f / a->loss_scale, a->scale_by_rays_info, a->learning_rate, a->d_step, a->lr_decay, a->beta1,
               a->beta2, a->epsilon, a->weights,
               a->lp_dtype == 1 ? static_cast<_Float16*>(a->weights_lp) : nullptr, a->first_moments,
               a->second_moments, a->param_steps};
    return true;
}

static int hb_planned(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride, uint32_t dfeat_stride,
                      tn_hashgrid_t grid, const float* pos, const void* dfeat, float* dtable, const void* plan,
                      size_t plan_bytes, void* work, size_t work_bytes, const HbAdam& adam, hipStream_t stream) {
    TN_CHECK_ARG(pos && dfeat && dtable && plan && work && n_max <= stride && n_max <= dfeat_stride);
    TN_CHECK_ARG(dfeat_dtype == 0 || dfeat_dtype == 1);
    HbPlan p;
    TN_CHECK_ARG(n_max == 0 || hb_split_plan(grid, n_max, p));
    if (n_max == 0) return TN_OK;
    TN_CHECK_ARG(plan_bytes >= hb_plan_area(p, nullptr, nullptr) && work_bytes >= hb_work_area(p, nullptr, nullptr));
    TN_CHECK_ARG(((reinterpret_cast<uintptr_t>(plan) | reinterpret_cast<uintptr_t>(work)) & 255) == 0 &&
                 (reinterpret_cast<uintptr_t>(dtable) & 7) == 0);
    HbPlanArea pa;
    HbWorkArea wa;
    hb_plan_area(p, const_cast<void*>(plan), &pa);
    hb_work_area(p, work, &wa);
    if (dfeat_dtype == 0)
        hashgrid_backward_planned<float>(n_max, info, stride, dfeat_stride, grid, p, pos,
                                         static_cast<const float*>(dfeat), dtable, pa, wa, stream, adam);
    else
        hashgrid_backward_planned<_Float16>(n_max, info, stride, dfeat_stride, grid, p, pos,
                                            static_cast<const _Float16*>(dfeat), dtable, pa, wa, stream, adam);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_hashgrid_backward_planned(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
                                            uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos,
                                            const void* dfeat, float* dtable, const void* plan, size_t plan_bytes,
                                            void* work, size_t work_bytes, hipStream_t stream) {
    return hb_planned(dfeat_dtype, n_max, info, stride, dfeat_stride, grid, pos, dfeat, dtable, plan, plan_bytes,
                      work, work_bytes, HbAdam{}, stream);
}

// tn_hashgrid_backward_planned with NGP Adam applied to the hash table where
// each entry's gradient is formed, in place of tn_adam_step over the table's
// range (same parameters, moments and steps bit for bit); dtable is a
// zero-filled gradient scratch that only non-finite or oversized split bins
// pass through, left zero again.
extern "C" int tn_hashgrid_backward_planned_adam(int dfeat_dtype, uint32_t n_max, const uint32_t* info,
                                                 uint32_t stride, uint32_t dfeat_stride, tn_hashgrid_t grid,
                                                 const float* pos, const void* dfeat, float* dtable, const void* plan,
                                                 size_t plan_bytes, void* work, size_t work_bytes,
                                                 const tn_adam_t* adam, hipStream_t stream) {
    HbAdam a;
    TN_CHECK_ARG(hb_adam_from(adam, a));
    return hb_planned(dfeat_dtype, n_max, info, stride, dfeat_stride, grid, pos, dfeat, dtable, plan, plan_bytes,
                      work, work_bytes, a, stream);
}

// The gradient passes of tn_hashgrid_backward_planned in parts, for a caller
// whose feature gradients of a batch become ready in sample ranges: the
// records of samples [sample_begin, sample_end) of every level are scattered
// as soon as their gradients exist (ranges are whole count blocks of
// tn_hashgrid_backward_part_samples samples each; the end may be n_max), and
// tn_hashgrid_backward_accumulate sums them once every range of the batch is
// scattered. Same result as tn_hashgrid_backward_planned bit for bit.
extern "C" uint32_t tn_hashgrid_backward_part_samples(uint32_t n_max, tn_hashgrid_t grid) {
    HbPlan p;
    return hb_split_plan(grid, n_max, p) ? p.chunk_blocks * 256u : 0u;
}

extern "C" int tn_hashgrid_backward_scatter(int dfeat_dtype, uint32_t n_max, const uint32_t* info, uint32_t stride,
                                            uint32_t dfeat_stride, tn_hashgrid_t grid, const float* pos,
                                            const void* dfeat, const void* plan, size_t plan_bytes, void* work,
                                            size_t work_bytes, uint32_t sample_begin, uint32_t sample_end,
                                            hipStream_t stream) {
    TN_CHECK_ARG(pos && dfeat && plan && work && n_max <= stride && n_max <= dfeat_stride);
    TN_CHECK_ARG(dfeat_dtype == 0 || dfeat_dtype == 1);
    HbPlan p;
    TN_CHECK_ARG(n_max == 0 || hb_split_plan(grid, n_max, p));
    if (n_max == 0) return TN_OK;
    const uint32_t part = p.chunk_blocks * 256u;
    TN_CHECK_ARG(sample_begin <= sample_end && sample_begin % part == 0 &&
                 (sample_end % part == 0 || sample_end >= n_max));
    TN_CHECK_ARG(plan_bytes >= hb_plan_area(p, nullptr, nullptr) && work_bytes >= hb_work_area(p, nullptr, nullptr));
    TN_CHECK_ARG(((reinterpret_cast<uintptr_t>(plan) | reinterpret_cast<uintptr_t>(work)) & 255) == 0);
    const uint32_t j0 = std::min(sample_begin / part, p.level_blocks);
    // the last part runs every remaining block: blocks past the batch still
    // write their (zero) bin maxima, which the sums read
    const uint32_t j1 = sample_end >= n_max ? p.level_blocks : std::min(sample_end / part, p.level_blocks);
    HbPlanArea pa;
    HbWorkArea wa;
    hb_plan_area(p, const_cast<void*>(plan), &pa);
    hb_work_area(p, work, &wa);
    const uint32_t L = (uint32_t)grid.n_levels;
    if (dfeat_dtype == 0)
        hb_scatter<float>(n_max, info, stride, dfeat_stride, grid, p, pos, static_cast<const float*>(dfeat), pa, wa,
                          0, L, j0, j1, stream);
    else
        hb_scatter<_Float16>(n_max, info, stride, dfeat_stride, grid, p, pos, static_cast<const _Float16*>(dfeat),
                             pa, wa, 0, L, j0, j1, stream);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

static int hb_accumulate_all(uint32_t n_max, tn_hashgrid_t grid, float* dtable, const void* plan, size_t plan_bytes,
                             void* work, size_t work_bytes, const HbAdam& adam, hipStream_t stream) {
    TN_CHECK_ARG(dtable && plan && work);
    HbPlan p;
    TN_CHECK_ARG(n_max == 0 || hb_split_plan(grid, n_max, p));
    if (n_max == 0) return TN_OK;
    TN_CHECK_ARG(plan_bytes >= hb_plan_area(p, nullptr, nullptr) && work_bytes >= hb_work_area(p, nullptr, nullptr));
    TN_CHECK_ARG(((reinterpret_cast<uintptr_t>(plan) | reinterpret_cast<uintptr_t>(work)) & 255) == 0 &&
                 (reinterpret_cast<uintptr_t>(dtable) & 7) == 0);
    HbPlanArea pa;
    HbWorkArea wa;
    hb_plan_area(p, const_cast<void*>(plan), &pa);
    hb_work_area(p, work, &wa);
    hb_accumulate(n_max, grid, p, dtable, pa, wa, 0, (uint32_t)grid.n_levels, adam, stream);
    hb_finish(grid, p, dtable, pa, wa, adam, stream);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_hashgrid_backward_accumulate(uint32_t n_max, tn_hashgrid_t grid, float* dtable, const void* plan,
                                               size_t plan_bytes, void* work, size_t work_bytes,
                                               hipStream_t stream) {
    return hb_accumulate_all(n_max, grid, dtable, plan, plan_bytes, work, work_bytes, HbAdam{}, stream);
}

// tn_hashgrid_backward_accumulate with Adam on the hash table, as
// tn_hashgrid_backward_planned_adam
extern "C" int tn_hashgrid_backward_accumulate_adam(uint32_t n_max, tn_hashgrid_t grid, float* dtable,
                                                    const void* plan, size_t plan_bytes, void* work,
                                                    size_t work_bytes, const tn_adam_t* adam, hipStream_t stream) {
    HbAdam a;
    TN_CHECK_ARG(hb_adam_from(adam, a));
    return hb_accumulate_all(n_max, grid, dtable, plan, plan_bytes, work, work_bytes, a, stream);
}
