// Wave64 machinery for per-ray work over compacted samples.
//
// Samples of ray r occupy [offset[r], offset[r] + steps[r]) and rays are laid
// out back to back. A wave owns the rays whose first sample falls in its
// window [w*S, (w+1)*S) (the last wave also owns any zero-length rays at the
// end) and walks their samples in chunks of 64, one sample per lane. Inside a
// chunk, per-ray prefix sums are segmented inclusive scans keyed by ray
// index; a ray that continues into the next chunk carries its running sums.
// Sums in fp32 are compensated (TwoSum, Knuth TAOCP vol. 2 4.2.2) so a
// 1024-sample ray accumulates O(log n) rather than O(n) rounding error.
#pragma once

#include "common.hpp"

#include <climits>

namespace tn {

// ---------------------------------------------------------------------------
// compensated fp32 pair
// ---------------------------------------------------------------------------
struct CSum {
    float hi, lo;
};

TN_DEVICE CSum csum_add(CSum a, CSum b) {
    const float s = a.hi + b.hi;
    const float bp = s - a.hi;
    const float e = (a.hi - (s - bp)) + (b.hi - bp);
    CSum r;
    r.hi = s;
    r.lo = a.lo + b.lo + e;
    return r;
}

TN_DEVICE float csum_val(CSum a) { return a.hi + a.lo; }

// Segmented inclusive scan across the wave, segments keyed by `key` (lanes
// with equal key are contiguous), on DPP lane moves: rows of 16 first
// (row_shr 1, 2, 4, 8), then the last lane of rows 0 and 2 into rows 1 and 3
// (row_bcast:15) and the last lane of row 1 into rows 2 and 3
// (row_bcast:31). A lane combines the moved value only when the source lane
// has its key; sources outside the row and unselected rows read key
// INT_MIN, which no segment uses.
template <int CTRL, int ROWS>
TN_DEVICE int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
TN_DEVICE float dpp_f(float v) {
    return __int_as_float(dpp_i<CTRL, ROWS>(0, __float_as_int(v)));
}
template <int CTRL, int ROWS>
TN_DEVICE double dpp_d(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)dpp_i<CTRL, ROWS>(0, (int)(uint32_t)b);
    const uint32_t hi = (uint32_t)dpp_i<CTRL, ROWS>(0, (int)(uint32_t)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int CTRL, int ROWS>
TN_DEVICE void seg_step(CSum& v, int key) {
    const bool take = dpp_i<CTRL, ROWS>(INT_MIN, key) == key;
    const float oh = dpp_f<CTRL, ROWS>(v.hi), ol = dpp_f<CTRL, ROWS>(v.lo);
    if (take) v = csum_add(v, CSum{oh, ol});
}
template <int CTRL, int ROWS>
TN_DEVICE void seg_step(double& v, int key) {
    const bool take = dpp_i<CTRL, ROWS>(INT_MIN, key) == key;
    const double o = dpp_d<CTRL, ROWS>(v);
    if (take) v += o;
}

template <typename V>
TN_DEVICE V seg_scan(V v, int key) {
    seg_step<0x111, 0xF>(v, key);
    seg_step<0x112, 0xF>(v, key);
    seg_step<0x114, 0xF>(v, key);
    seg_step<0x118, 0xF>(v, key);
    seg_step<0x142, 0xA>(v, key);
    seg_step<0x143, 0xC>(v, key);
    return v;
}

// inclusive max-scan of a non-negative-or--1 int across the wave (DPP)
TN_DEVICE int max_scan(int v) {
    v = max(v, dpp_i<0x111, 0xF>(-1, v));
    v = max(v, dpp_i<0x112, 0xF>(-1, v));
    v = max(v, dpp_i<0x114, 0xF>(-1, v));
    v = max(v, dpp_i<0x118, 0xF>(-1, v));
    v = max(v, dpp_i<0x142, 0xA>(-1, v));
    v = max(v, dpp_i<0x143, 0xC>(-1, v));
    return v;
}

// Segmented inclusive scan from the high lanes down (suffix sums).
TN_DEVICE CSum seg_scan_rev(CSum v, int key) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const float oh = __shfl_down(v.hi, d, 64);
        const float ol = __shfl_down(v.lo, d, 64);
        const int ok = __shfl_down(key, d, 64);
        if (lane + d < 64 && ok == key) v = csum_add(v, CSum{oh, ol});
    }
    return v;
}

TN_DEVICE double seg_scan_rev(double v, int key) {
    const int lane = (int)__lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_down(v, d, 64);
        const int ok = __shfl_down(key, d, 64);
        if (lane + d < 64 && ok == key) v += o;
    }
    return v;
}

// accumulator abstraction: fp32 compensated or plain fp64
template <typename A>
struct Acc;

template <>
struct Acc<float> {
    CSum v;
    TN_DEVICE static Acc zero() { return Acc{CSum{0.0f, 0.0f}}; }
    TN_DEVICE static Acc of(float x) { return Acc{CSum{x, 0.0f}}; }
    TN_DEVICE float val() const { return csum_val(v); }
    TN_DEVICE Acc operator+(const Acc& o) const { return Acc{csum_add(v, o.v)}; }
    TN_DEVICE Acc scan(int key) const { return Acc{seg_scan(v, key)}; }
    TN_DEVICE Acc scan_rev(int key) const { return Acc{seg_scan_rev(v, key)}; }
    TN_DEVICE Acc bcast(int src) const {
        return Acc{CSum{__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.hi), src)),
                        __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v.lo), src))}};
    }
};

template <>
struct Acc<double> {
    double v;
    TN_DEVICE static Acc zero() { return Acc{0.0}; }
    TN_DEVICE static Acc of(double x) { return Acc{x}; }
    TN_DEVICE double val() const { return v; }
    TN_DEVICE Acc operator+(const Acc& o) const { return Acc{v + o.v}; }
    TN_DEVICE Acc scan(int key) const { return Acc{seg_scan(v, key)}; }
    TN_DEVICE Acc scan_rev(int key) const { return Acc{seg_scan_rev(v, key)}; }
    TN_DEVICE Acc bcast(int src) const {
        const uint64_t b = (uint64_t)__double_as_longlong(v);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
        return Acc{__longlong_as_double((long long)(((uint64_t)hi << 32) | lo))};
    }
};

template <typename R>
TN_DEVICE R r_exp(R x);
template <>
TN_DEVICE float r_exp<float>(float x) { return __expf(x); }
template <>
TN_DEVICE double r_exp<double>(double x) { return exp(x); }

// alpha = 1 - exp(-tau) without the cancellation of the subtraction when
// tau is small (fine steps through thin media)
template <typename R>
TN_DEVICE R r_alpha(R tau);
template <>
TN_DEVICE float r_alpha<float>(float tau) { return -expm1f(-tau); }
template <>
TN_DEVICE double r_alpha<double>(double tau) { return -expm1(-tau); }

// ---------------------------------------------------------------------------
// first index r in [0, n) with offs[r] >= x (n if none), 64-ary search
// ---------------------------------------------------------------------------
TN_DEVICE uint32_t wave_lower_bound(const uint32_t* __restrict__ offs, uint32_t n, uint32_t x) {
    const uint32_t lane = __lane_id();
    uint32_t lo = 0, hi = n;  // answer in [lo, hi]
    while (lo < hi) {
        const uint32_t span = hi - lo;
        const uint32_t step = (span + 63) / 64;
        const uint32_t p = lo + lane * step;
        const bool pred = p < hi && offs[p] >= x;
        const bool inrange = p < hi;
        const uint64_t ball = __ballot(pred);
        const uint64_t inr = __ballot(inrange);
        if (ball == 0) {
            const uint32_t last = 63u - (uint32_t)__builtin_clzll(inr);
            lo = lo + last * step + 1;
        } else {
            const uint32_t f = (uint32_t)__builtin_ctzll(ball);
            if (f == 0) return lo;
            hi = lo + f * step;
            lo = lo + (f - 1) * step + 1;
        }
    }
    return lo;
}

// Ray range and sample window of wave `w` with window S.
struct WaveRange {
    uint32_t r0, r1, s0, s1;
    bool active;
};

TN_DEVICE WaveRange wave_range(uint32_t w, uint32_t S, uint32_t n_rays, uint32_t n_samples,
                               const uint32_t* __restrict__ offs) {
    WaveRange wr;
    const uint32_t n_waves = n_samples == 0 ? 1 : (n_samples + S - 1) / S;
    wr.active = w < n_waves && n_rays > 0;
    if (!wr.active) return wr;
    wr.r0 = w == 0 ? 0 : wave_lower_bound(offs, n_rays, w * S);
    wr.r1 = (w + 1 == n_waves) ? n_rays : wave_lower_bound(offs, n_rays, (w + 1) * S);
    wr.s0 = wr.r0 < n_rays ? min(offs[wr.r0], n_samples) : n_samples;
    wr.s1 = wr.r1 < n_rays ? min(offs[wr.r1], n_samples) : n_samples;
    return wr;
}

// ---------------------------------------------------------------------------
// Walking a window's rays chunk by chunk. Rays with index >= r_scan start at
// or after the current chunk; carry_ray (or -1) continues from the previous
// chunk and ends at sample carry_end. Each lane caches the offset and step
// count of ray r_cache + lane, reloaded only when r_scan moves (when rays
// start), so a chunk inside a long ray issues no loads.
// ---------------------------------------------------------------------------
struct RayWalk {
    uint32_t r_scan;
    uint32_t r_cache = UINT_MAX;
    uint32_t o = 0, n = 0;  // per lane
    int carry_ray = -1;
    int carry_end = 0;
};

// the ray of a lane's sample and the end (one past the last sample) of it
struct LaneRay {
    int ray;
    int end;
};

// Ray of each lane in the chunk [c, c + 64). Heads of rays with samples are
// scattered into `slot` (a 128-entry LDS row owned by this wave: ray indices,
// then ray ends); zero-length rays met on the way are reported through
// `on_empty(ray)` (called by the lane that holds the ray). Lanes before the
// first head take the carried ray. Sample offsets are below 2^31.
template <typename OnEmpty>
TN_DEVICE LaneRay chunk_rays(uint32_t c, uint32_t r1, RayWalk& wk, const uint32_t* __restrict__ offs,
                             const uint32_t* __restrict__ steps, int* slot, OnEmpty&& on_empty) {
    const uint32_t lane = __lane_id();
    slot[lane] = -1;
    slot[64 + lane] = -1;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    while (true) {
        const uint32_t k = wk.r_scan + lane;
        if (wk.r_scan != wk.r_cache) {
            wk.o = k < r1 ? offs[k] : UINT_MAX;
            wk.n = k < r1 ? steps[k] : 0u;
            wk.r_cache = wk.r_scan;
        }
        const bool inchunk = k < r1 && wk.o < c + 64u;
        if (inchunk) {
            if (wk.n > 0) {
                slot[wk.o - c] = (int)k;
                slot[64 + wk.o - c] = (int)(wk.o + wk.n);
            } else {
                on_empty(k);
            }
        }
        const uint32_t cnt = (uint32_t)__builtin_popcountll(__ballot(inchunk));
        wk.r_scan += cnt;
        if (cnt < 64) break;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // propagate the most recent head to the following lanes (max-scans: ray
    // indices and ray ends increase along the chunk)
    const int head = max_scan(slot[lane]);
    const int end = max_scan(slot[64 + lane]);
    __builtin_amdgcn_wave_barrier();
    return head < 0 ? LaneRay{wk.carry_ray, wk.carry_end} : LaneRay{head, end};
}

// ---------------------------------------------------------------------------
// Chunk-parallel passes over a batch cut into chunks of 64 samples, one per
// wave, with per-ray chaining of the chunk sums between them.
// ---------------------------------------------------------------------------
// Walk state at chunk start c: the ray holding sample c is carried into the
// chunk when it started earlier; otherwise the walk starts at the first ray
// beginning at c (zero-length rays there included).
TN_DEVICE RayWalk chunk_walk(uint32_t c, uint32_t n_rays, const uint32_t* __restrict__ offs,
                             const uint32_t* __restrict__ steps) {
    RayWalk wk;
    const uint32_t r_next = wave_lower_bound(offs, n_rays, c + 1);  // offs[0] = 0 <= c, so r_next >= 1
    const uint32_t rf = r_next - 1;
    const uint32_t of = offs[rf];
    if (of < c) {
        wk.carry_ray = (int)rf;
        wk.carry_end = (int)(of + steps[rf]);
        wk.r_scan = r_next;
    } else {
        wk.r_scan = wave_lower_bound(offs, n_rays, c);
    }
    return wk;
}

template <typename R>
TN_DEVICE Acc<R> acc_scaled(const Acc<R>& a, R k);
template <>
TN_DEVICE Acc<float> acc_scaled<float>(const Acc<float>& a, float k) {
    return Acc<float>{CSum{a.v.hi * k, a.v.lo * k}};
}
template <>
TN_DEVICE Acc<double> acc_scaled<double>(const Acc<double>& a, double k) {
    return Acc<double>{a.v * k};
}

// exclusive scan of an accumulator across the wave (lane 0 gets zero)
template <typename R>
TN_DEVICE Acc<R> acc_exclusive(const Acc<R>& v) {
    const Acc<R> inc = v.scan(0);
    Acc<R> ex;
    if constexpr (sizeof(R) == 4) {
        ex = Acc<R>{CSum{__shfl_up(inc.v.hi, 1, 64), __shfl_up(inc.v.lo, 1, 64)}};
    } else {
        ex = Acc<R>{__shfl_up(inc.v, 1, 64)};
    }
    return __lane_id() == 0 ? Acc<R>::zero() : ex;
}

// value of lane 63
TN_DEVICE float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
TN_DEVICE int lane63(int v) { return __builtin_amdgcn_readlane(v, 63); }
TN_DEVICE double lane63(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

}  // namespace tn
