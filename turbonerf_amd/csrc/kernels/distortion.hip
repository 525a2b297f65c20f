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
// reference's running-total adjoint rewritten with three lane-parallel passes
// over each wave's rays: totals (and the loss), sum_j w_j g_j, gradients.
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

// scratch rows per ray: W total, WM total, sum_j w_j g_j
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kDistWaves)
distortion_kernel(const uint32_t* __restrict__ info, uint32_t S, float lambda, const uint32_t* __restrict__ steps,
                  const uint32_t* __restrict__ offs, const TIn* __restrict__ density, const float* __restrict__ mn,
                  const float* __restrict__ dtn, float* __restrict__ ray_loss, float* __restrict__ dL_ds,
                  R* __restrict__ scratch, uint32_t SB) {
    __shared__ int slots[kDistWaves][128];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    const uint32_t w = blockIdx.x * kDistWaves + wid;
    const uint32_t n_rays = info[0], n_samples = info[1];
    const WaveRange wr = wave_range(w, S, n_rays, n_samples, offs);
    if (!wr.active) return;
    int* slot = slots[wid];

    for (int pass = 0; pass < 3; ++pass) {
        RayWalk wk;
        wk.r_scan = wr.r0;
        Acc<R> cP = Acc<R>::zero(), cW = Acc<R>::zero(), cWM = Acc<R>::zero(), cL = Acc<R>::zero();
        auto empty_ray = [&](uint32_t r) {
            if (pass == 0) ray_loss[r] = 0.0f;
        };
        for (uint32_t c = wr.s0; c < wr.s1; c += 64) {
            const uint64_t s = (uint64_t)c + lane;
            const bool valid = s < wr.s1;
            const int prev = wk.carry_ray;
            const LaneRay lr = chunk_rays(c, wr.r1, wk, offs, steps, slot, empty_ray);
            const int ray = valid ? lr.ray : -2 - (int)lane;
            const DistState<R> st = dist_state<R>(valid, ray, prev, s, density, mn, dtn, cP, cW, cWM);
            const uint32_t rr = valid ? (uint32_t)ray : 0u;
            const bool tail = valid && s + 1 == (uint64_t)lr.end;
            const R kf = valid ? (R)lambda / (R)steps[rr] : (R)0;
            R v = 0;
            if (pass == 0) {
                // loss terms
                const R W_ex = st.W_incl - st.w, WM_ex = st.WM_incl - st.w * st.m;
                v = st.d * st.w * st.w / (R)3 + (R)2 * st.w * (st.m * W_ex - WM_ex);
            } else {
                const R Wt = valid ? scratch[rr] : (R)0, WMt = valid ? scratch[rr + SB] : (R)0;
                const R W_ex = st.W_incl - st.w, WM_ex = st.WM_incl - st.w * st.m;
                const R g = kf * ((R)2 / (R)3 * st.d * st.w + (R)2 * (st.m * W_ex - WM_ex) +
                                  (R)2 * ((WMt - st.WM_incl) - st.m * (Wt - st.W_incl)));
                v = st.w * g;
                if (pass == 2) {
                    // suffix sum_{j>i} w_j g_j = total - inclusive prefix
                    Acc<R> Q = Acc<R>::of(v).scan(ray);
                    if (ray == prev) Q = Q + cL;
                    if (valid) {
                        const R suffix = scratch[rr + 2 * SB] - Q.val();
                        dL_ds[s] = (float)(st.d * (st.T_next * g - suffix) * st.sigma);
                    }
                    cL = Q.bcast(63);
                }
            }
            if (pass < 2) {
                Acc<R> L = Acc<R>::of(v).scan(ray);
                if (ray == prev) L = L + cL;
                if (tail) {
                    if (pass == 0) {
                        ray_loss[rr] = (float)(kf * L.val());
                        scratch[rr] = st.W_incl;
                        scratch[rr + SB] = st.WM_incl;
                    } else {
                        scratch[rr + 2 * SB] = L.val();
                    }
                }
                cL = L.bcast(63);
            }
            const bool tail63 = lane63((int)tail) != 0;
            wk.carry_ray = lane63(ray);
            wk.carry_end = lane63(lr.end);
            if (tail63 || !lane63((int)valid)) wk.carry_ray = -1;
        }
        if (pass == 0)
            for (uint32_t r = wk.r_scan + lane; r < wr.r1; r += 64) empty_ray(r);
        // the next pass reads totals this wave's tail lanes wrote
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_distortion_scratch_bytes(uint32_t n_rays_max) { return (size_t)3 * n_rays_max * sizeof(double); }

// dtype: 0 = fp32 density, 1 = fp16 density, 2 = fp32 with fp64 accumulation.
// Per-ray loss -> ray_loss[n_rays]; dL/ds -> dL_ds[sample_stride] (unscaled;
// the MLP backward applies the loss scale).
extern "C" int tn_distortion(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t sample_stride, float lambda,
                             const uint32_t* ray_steps, const uint32_t* ray_offset, const void* density,
                             const float* m_norm, const float* dt_norm, float* ray_loss, float* dL_ds, void* scratch,
                             size_t scratch_bytes, hipStream_t stream) {
    TN_CHECK_ARG(info && ray_steps && ray_offset && density && m_norm && dt_norm && ray_loss && dL_ds && scratch);
    TN_CHECK_ARG(n_rays_max > 0 && sample_stride > 0 && sample_stride < (1u << 31));
    if (scratch_bytes < tn_distortion_scratch_bytes(n_rays_max)) return TN_ERR_OUT_OF_MEMORY;
    const uint32_t S = 128;
    const uint32_t waves = (sample_stride + S - 1) / S + 1;
    const unsigned blocks = (waves + kDistWaves - 1) / kDistWaves;
    switch (dtype) {
        case 0:
            distortion_kernel<float, float><<<blocks, 64 * kDistWaves, 0, stream>>>(
                info, S, lambda, ray_steps, ray_offset, static_cast<const float*>(density), m_norm, dt_norm, ray_loss,
                dL_ds, static_cast<float*>(scratch), n_rays_max);
            break;
        case 1:
            distortion_kernel<_Float16, float><<<blocks, 64 * kDistWaves, 0, stream>>>(
                info, S, lambda, ray_steps, ray_offset, static_cast<const _Float16*>(density), m_norm, dt_norm,
                ray_loss, dL_ds, static_cast<float*>(scratch), n_rays_max);
            break;
        case 2:
            distortion_kernel<float, double><<<blocks, 64 * kDistWaves, 0, stream>>>(
                info, S, lambda, ray_steps, ray_offset, static_cast<const float*>(density), m_norm, dt_norm, ray_loss,
                dL_ds, static_cast<double*>(scratch), n_rays_max);
            break;
        default:
            return TN_ERR_INVALID_ARGUMENT;
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
}
