// mip-NeRF 360 distortion loss and its gradient, O(N) prefix-sum form
// (reference src/utils/training-kernels-fused.cuh:13-224, orchestration at
// src/core/nerf-network.cu:605-666). Per ray, with m the normalised sample
// midpoints, d the normalised steps and w_i = T_i (1 - exp(-sigma_i d_i)):
//     L = (lambda/N) [ (1/3) sum_i d_i w_i^2 + 2 sum_i w_i (m_i W_{i-1} - WM_{i-1}) ]
// where W, WM are prefix sums of w and w m. As in the reference, alpha uses
// the normalised step and there is no early termination.
//
// Gradient: for g_j = dL/dw_j,
//     g_j/k = (2/3) d_j w_j + 2 (m_j W_{j-1} - WM_{j-1}) + 2 ((WM - WM_j) - m_j (W - W_j))
//     dL/dsigma_i = d_i (T_{i+1} g_i - sum_{j>i} w_j g_j),   dL/ds_i = sigma_i dL/dsigma_i
// (k = lambda/N, unsubscripted W, WM are ray totals). This is the
// reference's running-total adjoint rewritten for chunk-parallel passes (the
// batch cut into chunks of 64 samples, one per wave), none of which walks a
// long ray serially:
//  (1) per chunk, segmented scans of sigma*d, w and w m from each segment's
//      start (transmittance 1 there): a record for the segments entering and
//      leaving the chunk, the totals W, WM of rays complete in the chunk;
//  (2) a wave per ray crossing chunks chains its records with lane scans
//      (a segment's w sums scale by the transmittance at its start): the
//      prefixes of sigma*d, W and WM at each chunk start and the totals;
//  (3) per chunk, the per-sample terms with those carry-ins: the loss term
//      v_i = d w^2/3 + 2 w (m W_{i-1} - WM_{i-1}) and w_i g_i, summed per
//      segment (the loss and sum w g of complete rays written here);
//  (4) a wave per ray crossing chunks: the loss, sum_j w_j g_j and its prefix
//      at each chunk start;
//  (5) per chunk, the gradient with every carry-in.
#include "segscan.hpp"

namespace tn {

constexpr int kDistWaves = 4;

template <typename T>
TN_DEVICE float ldv(const T* p, uint64_t i) { return (float)p[i]; }

template <typename R>
struct DistState {
    R sigma, d, m, T_next, w, W_incl, WM_incl;
};

// per-chunk recomputation: transmittance, weights and the prefix sums W, WM
template <typename R, typename TIn>
TN_DEVICE DistState<R> dist_state(bool valid, int ray, int carry_ray_in, uint64_t s, const TIn* __restrict__ density,
                                  const float* __restrict__ mn, const float* __restrict__ dtn, Acc<R>& cP, Acc<R>& cW,
                                  Acc<R>& cWM) {
    DistState<R> st;
    st.sigma = 0;
    st.d = 0;
    st.m = 0;
    if (valid) {
        st.sigma = r_exp<R>((R)fminf(ldv(density, s), 12.0f) - (R)1);
        st.d = (R)dtn[s];
        st.m = (R)mn[s];
    }
    const R tau = st.sigma * st.d;
    Acc<R> P = Acc<R>::of(tau).scan(ray);
    if (ray == carry_ray_in) P = P + cP;
    const R p_incl = P.val();
    const R T = r_exp<R>(-(p_incl - tau));
    st.T_next = r_exp<R>(-p_incl);
    st.w = valid ? T * r_alpha<R>(tau) : (R)0;
    Acc<R> W = Acc<R>::of(st.w).scan(ray);
    Acc<R> WM = Acc<R>::of(st.w * st.m).scan(ray);
    if (ray == carry_ray_in) {
        W = W + cW;
        WM = WM + cWM;
    }
    st.W_incl = W.val();
    st.WM_incl = WM.val();
    cP = P.bcast(63);
    cW = W.bcast(63);
    cWM = WM.bcast(63);
    return st;
}

// per chunk: two records (the segments of lanes 0 and 63) of the sums of
// sigma*d, w and w m, then of the sums of v and w g; carry-ins per chunk of
// the prefixes of sigma*d, W, WM and of w g
constexpr int kDistRec1 = 3, kDistRec2 = 2, kDistCarry = 4;

inline size_t dist_align(size_t x) { return (x + 255) & ~(size_t)255; }
inline uint32_t dist_chunks(uint32_t sample_stride) { return sample_stride / 64 + 1; }
inline size_t dist_rays_bytes(uint32_t n_rays) { return dist_align((size_t)3 * n_rays * 8); }
inline size_t dist_rec1_bytes(uint32_t ss) { return dist_align((size_t)dist_chunks(ss) * 2 * kDistRec1 * 8); }
inline size_t dist_rec2_bytes(uint32_t ss) { return dist_align((size_t)dist_chunks(ss) * 2 * kDistRec2 * 8); }
inline size_t dist_carry_bytes(uint32_t ss) { return dist_align((size_t)dist_chunks(ss) * kDistCarry * 8); }

// the segments of lane 0 and lane 63 leave their sums in the chunk's records
template <typename R, int N>
TN_DEVICE void put_seg_recs(Acc<R>* __restrict__ rq, bool valid, int ray, const Acc<R> (&v)[N]) {
    const uint32_t lane = __lane_id();
    const bool seg_last = lane == 63 || __shfl_down(ray, 1, 64) != ray;
    if (valid && seg_last && ray == __builtin_amdgcn_readlane(ray, 0))
        for (int k = 0; k < N; ++k) rq[k] = v[k];
    if (valid && lane == 63)
        for (int k = 0; k < N; ++k) rq[N + k] = v[k];
}

// (1) chunk sums of sigma*d, w and w m from each segment's start
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
dist_chunk_kernel(const uint32_t* __restrict__ info, const uint32_t* __restrict__ steps,
                  const uint32_t* __restrict__ offs, const TIn* __restrict__ density, const float* __restrict__ mn,
                  const float* __restrict__ dtn, R* __restrict__ rays, uint32_t SB, Acc<R>* __restrict__ rec) {
    __shared__ int slots[kDistWaves][128];
    const uint32_t lane = __lane_id();
    const uint32_t q = blockIdx.x * kDistWaves + (threadIdx.x >> 6);
    const uint32_t n_rays = info[0], n_samples = info[1];
    if (n_rays == 0 || q >= (n_samples + 63) / 64) return;
    const uint32_t c = 64 * q;
    RayWalk wk = chunk_walk(c, n_rays, offs, steps);
    const int carried = wk.carry_ray;
    auto no_op = [](uint32_t) {};
    const LaneRay lr = chunk_rays(c, n_rays, wk, offs, steps, slots[threadIdx.x >> 6], no_op);
    const uint64_t s = (uint64_t)c + lane;
    const bool valid = s < n_samples;
    const int ray = valid ? lr.ray : -2 - (int)lane;
    Acc<R> zP = Acc<R>::zero(), zW = Acc<R>::zero(), zWM = Acc<R>::zero();
    const DistState<R> st = dist_state<R>(valid, ray, -1, valid ? s : 0, density, mn, dtn, zP, zW, zWM);
    const Acc<R> P = Acc<R>::of(st.sigma * st.d).scan(ray);
    const Acc<R> W = Acc<R>::of(st.w).scan(ray), WM = Acc<R>::of(st.w * st.m).scan(ray);
    if (valid && ray != carried && s + 1 == (uint64_t)lr.end) {  // a ray complete in this chunk
        rays[ray] = W.val();
        rays[ray + SB] = WM.val();
    }
    const Acc<R> v[kDistRec1] = {P, W, WM};
    put_seg_recs<R, kDistRec1>(rec + (uint64_t)q * 2 * kDistRec1, valid, ray, v);
}

// the records of ray chunk k (qs: the segment of lane 63, later: of lane 0)
template <typename R>
TN_DEVICE const Acc<R>* dist_rec(const Acc<R>* rec, int per, uint32_t k, uint32_t qs) {
    return rec + ((uint64_t)k * 2 + (k == qs ? 1u : 0u)) * per;
}

// (2) a wave per ray crossing chunks: prefixes at the chunk starts, totals
template <typename R>
TN_DEVICE void dist_prefix_one_ray(uint32_t r, const uint32_t* __restrict__ info, const uint32_t* __restrict__ steps,
                   const uint32_t* __restrict__ offs, float* __restrict__ ray_loss, R* __restrict__ rays, uint32_t SB,
                   const Acc<R>* __restrict__ rec, Acc<R>* __restrict__ carry) {
    const uint32_t lane = __lane_id();
    if (r >= info[0]) return;
    const uint32_t o = offs[r], n = steps[r];
    if (n == 0) {
        if (lane == 0) ray_loss[r] = 0.0f;
        return;
    }
    const uint32_t qs = o / 64, qe = (o + n - 1) / 64;
    if (qs == qe) return;  // complete in one chunk
    Acc<R> P0 = Acc<R>::zero(), W0 = Acc<R>::zero(), WM0 = Acc<R>::zero();
    for (uint32_t k0 = qs; k0 <= qe; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool act = k <= qe;
        const Acc<R>* rk = dist_rec<R>(rec, kDistRec1, act ? k : qe, qs);
        const Acc<R> Pk = act ? rk[0] : Acc<R>::zero();
        const Acc<R> Pin = P0 + acc_exclusive<R>(Pk);
        const R T = r_exp<R>(-Pin.val());
        const Acc<R> Wk = act ? acc_scaled<R>(rk[1], T) : Acc<R>::zero();
        const Acc<R> WMk = act ? acc_scaled<R>(rk[2], T) : Acc<R>::zero();
        const Acc<R> Win = W0 + acc_exclusive<R>(Wk), WMin = WM0 + acc_exclusive<R>(WMk);
        if (act && k > qs) {
            Acc<R>* cq = carry + (uint64_t)k * kDistCarry;
            cq[0] = Pin;
            cq[1] = Win;
            cq[2] = WMin;
        }
        P0 = (Pin + Pk).bcast(63);
        W0 = (Win + Wk).bcast(63);
        WM0 = (WMin + WMk).bcast(63);
    }
    if (lane == 0) {
        rays[r] = W0.val();
        rays[r + SB] = WM0.val();
    }
}

// rays r = wave, wave + the grid's waves, ...: the grid is capped (most of a
// training batch's ray capacity is unused), so idle workgroups are not launched
template <typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
dist_prefix_kernel(const uint32_t* __restrict__ info, const uint32_t* __restrict__ steps,
                   const uint32_t* __restrict__ offs, float* __restrict__ ray_loss, R* __restrict__ rays, uint32_t SB,
                   const Acc<R>* __restrict__ rec, Acc<R>* __restrict__ carry) {
    const uint32_t n_rays = info[0];
    for (uint32_t r = blockIdx.x * kDistWaves + (threadIdx.x >> 6); r < n_rays; r += gridDim.x * kDistWaves)
        dist_prefix_one_ray<R>(r, info, steps, offs, ray_loss, rays, SB, rec, carry);
}

// per-sample terms of pass (3) and (5): state with the carry-ins, the loss
// term v and g = dL/dw
template <typename R>
struct DistTerms {
    DistState<R> st;
    R v, g;
};

template <typename TIn, typename R>
TN_DEVICE DistTerms<R> dist_terms(bool valid, int ray, int carried, uint64_t s, float lambda,
                                  const uint32_t* __restrict__ steps, const TIn* __restrict__ density,
                                  const float* __restrict__ mn, const float* __restrict__ dtn,
                                  const R* __restrict__ rays, uint32_t SB, const Acc<R>* __restrict__ cq) {
    Acc<R> cP = Acc<R>::zero(), cW = Acc<R>::zero(), cWM = Acc<R>::zero();
    if (carried >= 0) {
        cP = cq[0];
        cW = cq[1];
        cWM = cq[2];
    }
    DistTerms<R> t;
    t.st = dist_state<R>(valid, ray, carried, valid ? s : 0, density, mn, dtn, cP, cW, cWM);
    const DistState<R>& st = t.st;
    const R W_ex = st.W_incl - st.w, WM_ex = st.WM_incl - st.w * st.m;
    t.v = st.d * st.w * st.w / (R)3 + (R)2 * st.w * (st.m * W_ex - WM_ex);
    t.g = 0;
    if (valid) {
        const uint32_t rr = (uint32_t)ray;
        const R kf = (R)lambda / (R)steps[rr];
        const R Wt = rays[rr], WMt = rays[rr + SB];
        t.g = kf * ((R)2 / (R)3 * st.d * st.w + (R)2 * (st.m * W_ex - WM_ex) +
                    (R)2 * ((WMt - st.WM_incl) - st.m * (Wt - st.W_incl)));
    }
    return t;
}

// (3) per chunk: the segment sums of v and w g; complete rays' loss and sum w g
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
dist_terms_kernel(const uint32_t* __restrict__ info, float lambda, const uint32_t* __restrict__ steps,
                  const uint32_t* __restrict__ offs, const TIn* __restrict__ density, const float* __restrict__ mn,
                  const float* __restrict__ dtn, float* __restrict__ ray_loss, R* __restrict__ rays, uint32_t SB,
                  const Acc<R>* __restrict__ carry, Acc<R>* __restrict__ rec2) {
    __shared__ int slots[kDistWaves][128];
    const uint32_t lane = __lane_id();
    const uint32_t q = blockIdx.x * kDistWaves + (threadIdx.x >> 6);
    const uint32_t n_rays = info[0], n_samples = info[1];
    if (n_rays == 0 || q >= (n_samples + 63) / 64) return;
    const uint32_t c = 64 * q;
    RayWalk wk = chunk_walk(c, n_rays, offs, steps);
    const int carried = wk.carry_ray;
    auto no_op = [](uint32_t) {};
    const LaneRay lr = chunk_rays(c, n_rays, wk, offs, steps, slots[threadIdx.x >> 6], no_op);
    const uint64_t s = (uint64_t)c + lane;
    const bool valid = s < n_samples;
    const int ray = valid ? lr.ray : -2 - (int)lane;
    const DistTerms<R> t = dist_terms<TIn, R>(valid, ray, carried, s, lambda, steps, density, mn, dtn, rays, SB,
                                              carry + (uint64_t)q * kDistCarry);
    const Acc<R> V = Acc<R>::of(t.v).scan(ray), Sg = Acc<R>::of(t.st.w * t.g).scan(ray);
    if (valid && ray != carried && s + 1 == (uint64_t)lr.end) {  // a ray complete in this chunk
        const uint32_t rr = (uint32_t)ray;
        ray_loss[rr] = (float)((R)lambda / (R)steps[rr] * V.val());
        rays[rr + 2 * SB] = Sg.val();
    }
    const Acc<R> v[kDistRec2] = {V, Sg};
    put_seg_recs<R, kDistRec2>(rec2 + (uint64_t)q * 2 * kDistRec2, valid, ray, v);
}

// (4) a wave per ray crossing chunks: loss, sum w g, its prefix at chunk starts
template <typename R>
TN_DEVICE void dist_totals_one_ray(uint32_t r, const uint32_t* __restrict__ info, float lambda, const uint32_t* __restrict__ steps,
                   const uint32_t* __restrict__ offs, float* __restrict__ ray_loss, R* __restrict__ rays, uint32_t SB,
                   const Acc<R>* __restrict__ rec2, Acc<R>* __restrict__ carry) {
    const uint32_t lane = __lane_id();
    if (r >= info[0]) return;
    const uint32_t o = offs[r], n = steps[r];
    if (n == 0) return;
    const uint32_t qs = o / 64, qe = (o + n - 1) / 64;
    if (qs == qe) return;
    Acc<R> V0 = Acc<R>::zero(), S0 = Acc<R>::zero();
    for (uint32_t k0 = qs; k0 <= qe; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool act = k <= qe;
        const Acc<R>* rk = dist_rec<R>(rec2, kDistRec2, act ? k : qe, qs);
        const Acc<R> Vk = act ? rk[0] : Acc<R>::zero(), Sk = act ? rk[1] : Acc<R>::zero();
        const Acc<R> Sin = S0 + acc_exclusive<R>(Sk);
        if (act && k > qs) carry[(uint64_t)k * kDistCarry + 3] = Sin;
        V0 = V0 + Vk.scan(0).bcast(63);
        S0 = (Sin + Sk).bcast(63);
    }
    if (lane == 0) {
        ray_loss[r] = (float)((R)lambda / (R)n * V0.val());
        rays[r + 2 * SB] = S0.val();
    }
}

// rays r = wave, wave + the grid's waves, ...: the grid is capped (most of a
// training batch's ray capacity is unused), so idle workgroups are not launched
template <typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
dist_totals_kernel(const uint32_t* __restrict__ info, float lambda, const uint32_t* __restrict__ steps,
                   const uint32_t* __restrict__ offs, float* __restrict__ ray_loss, R* __restrict__ rays, uint32_t SB,
                   const Acc<R>* __restrict__ rec2, Acc<R>* __restrict__ carry) {
    const uint32_t n_rays = info[0];
    for (uint32_t r = blockIdx.x * kDistWaves + (threadIdx.x >> 6); r < n_rays; r += gridDim.x * kDistWaves)
        dist_totals_one_ray<R>(r, info, lambda, steps, offs, ray_loss, rays, SB, rec2, carry);
}

// (5) per chunk: the gradient
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
dist_grad_kernel(const uint32_t* __restrict__ info, float lambda, const uint32_t* __restrict__ steps,
                 const uint32_t* __restrict__ offs, const TIn* __restrict__ density, const float* __restrict__ mn,
                 const float* __restrict__ dtn, float* __restrict__ dL_ds, const R* __restrict__ rays, uint32_t SB,
                 const Acc<R>* __restrict__ carry) {
    __shared__ int slots[kDistWaves][128];
    const uint32_t lane = __lane_id();
    const uint32_t q = blockIdx.x * kDistWaves + (threadIdx.x >> 6);
    const uint32_t n_rays = info[0], n_samples = info[1];
    if (n_rays == 0 || q >= (n_samples + 63) / 64) return;
    const uint32_t c = 64 * q;
    RayWalk wk = chunk_walk(c, n_rays, offs, steps);
    const int carried = wk.carry_ray;
    auto no_op = [](uint32_t) {};
    const LaneRay lr = chunk_rays(c, n_rays, wk, offs, steps, slots[threadIdx.x >> 6], no_op);
    const uint64_t s = (uint64_t)c + lane;
    const bool valid = s < n_samples;
    const int ray = valid ? lr.ray : -2 - (int)lane;
    const Acc<R>* cq = carry + (uint64_t)q * kDistCarry;
    const DistTerms<R> t = dist_terms<TIn, R>(valid, ray, carried, s, lambda, steps, density, mn, dtn, rays, SB, cq);
    Acc<R> Q = Acc<R>::of(t.st.w * t.g).scan(ray);
    if (ray == carried) Q = Q + cq[3];
    if (!valid) return;
    // suffix sum_{j>i} w_j g_j = total - inclusive prefix
    const R suffix = rays[(uint32_t)ray + 2 * SB] - Q.val();
    dL_ds[s] = (float)(t.st.d * (t.st.T_next * t.g - suffix) * t.st.sigma);
}

template <typename TIn, typename R>
void launch_distortion(const uint32_t* info, uint32_t n_rays_max, uint32_t sample_stride, float lambda,
                       const uint32_t* steps, const uint32_t* offs, const TIn* density, const float* mn,
                       const float* dtn, float* ray_loss, float* dL_ds, void* scratch, hipStream_t stream) {
    unsigned char* base = static_cast<unsigned char*>(scratch);
    R* rays = reinterpret_cast<R*>(base);
    size_t off = dist_rays_bytes(n_rays_max);
    Acc<R>* rec1 = reinterpret_cast<Acc<R>*>(base + off);
    off += dist_rec1_bytes(sample_stride);
    Acc<R>* rec2 = reinterpret_cast<Acc<R>*>(base + off);
    off += dist_rec2_bytes(sample_stride);
    Acc<R>* carry = reinterpret_cast<Acc<R>*>(base + off);
    const uint32_t SB = n_rays_max;
    const unsigned chunk_blocks = (dist_chunks(sample_stride) + kDistWaves - 1) / kDistWaves;
    const unsigned ray_blocks = ray_wave_blocks(n_rays_max, kDistWaves);
    const unsigned threads = 64 * kDistWaves;
    dist_chunk_kernel<TIn, R><<<chunk_blocks, threads, 0, stream>>>(info, steps, offs, density, mn, dtn, rays, SB,
                                                                     rec1);
    dist_prefix_kernel<R><<<ray_blocks, threads, 0, stream>>>(info, steps, offs, ray_loss, rays, SB, rec1, carry);
    dist_terms_kernel<TIn, R><<<chunk_blocks, threads, 0, stream>>>(info, lambda, steps, offs, density, mn, dtn,
                                                                     ray_loss, rays, SB, carry, rec2);
    dist_totals_kernel<R><<<ray_blocks, threads, 0, stream>>>(info, lambda, steps, offs, ray_loss, rays, SB, rec2,
                                                              carry);
    dist_grad_kernel<TIn, R><<<chunk_blocks, threads, 0, stream>>>(info, lambda, steps, offs, density, mn, dtn, dL_ds,
                                                                    rays, SB, carry);
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_distortion_scratch_bytes(uint32_t n_rays_max, uint32_t sample_stride) {
    return dist_rays_bytes(n_rays_max) + dist_rec1_bytes(sample_stride) + dist_rec2_bytes(sample_stride) +
           dist_carry_bytes(sample_stride);
}

// dtype: 0 = fp32 density, 1 = fp16 density, 2 = fp32 with fp64 accumulation.
// Per-ray loss -> ray_loss[n_rays]; dL/ds -> dL_ds[sample_stride] (unscaled;
// the MLP backward applies the loss scale).
extern "C" int tn_distortion(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t sample_stride, float lambda,
                             const uint32_t* ray_steps, const uint32_t* ray_offset, const void* density,
                             const float* m_norm, const float* dt_norm, float* ray_loss, float* dL_ds, void* scratch,
                             size_t scratch_bytes, hipStream_t stream) {
    TN_CHECK_ARG(info && ray_steps && ray_offset && density && m_norm && dt_norm && ray_loss && dL_ds && scratch);
    TN_CHECK_ARG(n_rays_max > 0 && sample_stride > 0 && sample_stride < (1u << 31));
    if (scratch_bytes < tn_distortion_scratch_bytes(n_rays_max, sample_stride)) return TN_ERR_OUT_OF_MEMORY;
    TN_CHECK_ARG((reinterpret_cast<uintptr_t>(scratch) & 255) == 0);
    switch (dtype) {
        case 0:
            launch_distortion<float, float>(info, n_rays_max, sample_stride, lambda, ray_steps, ray_offset,
                                            static_cast<const float*>(density), m_norm, dt_norm, ray_loss, dL_ds,
                                            scratch, stream);
            break;
        case 1:
            launch_distortion<_Float16, float>(info, n_rays_max, sample_stride, lambda, ray_steps, ray_offset,
                                               static_cast<const _Float16*>(density), m_norm, dt_norm, ray_loss,
                                               dL_ds, scratch, stream);
            break;
        case 2:
            launch_distortion<float, double>(info, n_rays_max, sample_stride, lambda, ray_steps, ray_offset,
                                             static_cast<const float*>(density), m_norm, dt_norm, ray_loss, dL_ds,
                                             scratch, stream);
            break;
        default:
            return TN_ERR_INVALID_ARGUMENT;
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
}
