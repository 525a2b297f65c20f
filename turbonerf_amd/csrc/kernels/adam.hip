// NGP Adam (reference src/core/adam-optimizer.cuh:48-96): per-parameter step
// counters, hash-table entries with a zero gradient are skipped entirely (no
// moment decay, no step), L2 regularisation on the network weights only, fp32
// master weights with a low-precision copy for the network kernels.
//
// Gradients are fp32 and are cleared as they are consumed, so the next
// step's scatter-adds start from zero without a separate memset over the
// whole table; entries whose gradient is zero are read once and skipped.
#include "common.hpp"

namespace tn {

template <typename TL>
TN_DEVICE void adam_one(uint32_t i, float g_raw, uint32_t n_grid, uint32_t n_net_end, float inv_loss_scale, float lr,
                        float beta1, float beta2, float eps, float l2, float* __restrict__ w, TL* __restrict__ w_lp,
                        float* __restrict__ m, float* __restrict__ v, uint32_t* __restrict__ steps) {
    float g = g_raw * inv_loss_scale;
    if (i < n_grid && g == 0.0f) return;
    float wf = w[i], mf = m[i], vf = v[i];
    if (i >= n_grid && i < n_net_end) g += l2 * wf;
    const uint32_t st = steps[i] + 1;
    steps[i] = st;
    adam_apply(g, adam_bias(st, beta1, beta2), lr, beta1, beta2, eps, wf, mf, vf);
    m[i] = mf;
    v[i] = vf;
    w[i] = wf;
    if (w_lp) w_lp[i] = (TL)wf;
}

template <typename TL>
struct LowPrecision4;
template <>
struct LowPrecision4<float> {
    TN_DEVICE static void store(float*, uint32_t, const float (&)[4]) {}
};
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
template <>
struct LowPrecision4<_Float16> {
    TN_DEVICE static void store(_Float16* p, uint32_t i4, const float (&w)[4]) {
        half4v h;
        h.x = (_Float16)w[0];
        h.y = (_Float16)w[1];
        h.z = (_Float16)w[2];
        h.w = (_Float16)w[3];
        *reinterpret_cast<half4v*>(p + i4) = h;
    }
};

// Streaming access to the fp32 master weights, moments and step counters,
// which the next kernels do not read (non-temporal, so they do not displace
// the gradient and low-precision weights from the caches).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef uint32_t nt_u4 __attribute__((ext_vector_type(4)));
TN_DEVICE float4 ld_nt(const float* p) {
    const nt_f4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f4*>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
TN_DEVICE uint4 ld_nt(const uint32_t* p) {
    const nt_u4 v = __builtin_nontemporal_load(reinterpret_cast<const nt_u4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}
TN_DEVICE void st_nt(float* p, const float (&a)[4]) {
    const nt_f4 v = {a[0], a[1], a[2], a[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_f4*>(p));
}
TN_DEVICE void st_nt(uint32_t* p, const uint32_t (&a)[4]) {
    const nt_u4 v = {a[0], a[1], a[2], a[3]};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_u4*>(p));
}

// One thread per 4 consecutive parameters with 16-byte loads and stores of
// the gradient, weights, moments and step counters (8 bytes for the fp16
// copy). The skip of zero-gradient hash entries is decided per 8 lanes, one
// 128-byte line of each fp32 array: a line without any gradient costs its
// gradient read only; otherwise all 32 parameters of the line are read and
// written back whole (skipped entries keep their values bit for bit), since
// partially written lines cost more than the bytes they save.

template <typename TL>
__global__ void adam_kernel(uint32_t n, uint32_t n_grid, uint32_t n_net_end, float inv_scale0,
                            const uint32_t* __restrict__ info, float lr0, const uint32_t* __restrict__ d_step,
                            float lr_decay, float beta1, float beta2, float eps, float l2, float* __restrict__ w,
                            TL* __restrict__ w_lp, float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, uint32_t* __restrict__ steps) {
    const uint32_t i4 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i4 >= n) return;
    // gradients scaled by 2^ceil(log2(rays)) as well come back exactly
    const float inv_loss_scale = info ? inv_scale0 / ray_pow2(info[0]) : inv_scale0;
    const float lr = step_lr(lr0, d_step, lr_decay);
    if (i4 + 3 < n) {
        const float4 g4 = *reinterpret_cast<const float4*>(grad + i4);
        const bool any = g4.x != 0.0f || g4.y != 0.0f || g4.z != 0.0f || g4.w != 0.0f;
        const uint64_t touched = __ballot(any);
        const bool line = ((touched >> (__lane_id() & 56u)) & 0xFFull) != 0ull;
        if (!line && i4 + 3 < n_grid) return;  // a line of untouched hash entries: its gradient read only
        const float4 w4 = ld_nt(w + i4);
        const float4 m4 = ld_nt(m + i4);
        const float4 v4 = ld_nt(v + i4);
        const uint4 s4 = ld_nt(steps + i4);
        const float gv[4] = {g4.x, g4.y, g4.z, g4.w};
        float wv[4] = {w4.x, w4.y, w4.z, w4.w}, mv[4] = {m4.x, m4.y, m4.z, m4.w}, vv[4] = {v4.x, v4.y, v4.z, v4.w};
        uint32_t sv[4] = {s4.x, s4.y, s4.z, s4.w};
        // the bias correction depends on the step count only; neighbours
        // mostly share it, so it is recomputed only when the count changes
        uint32_t c_st = 0;
        float c = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = i4 + k;
            float g = gv[k] * inv_loss_scale;
            if (i < n_grid && g == 0.0f) continue;  // skipped entry keeps its state
            if (i >= n_grid && i < n_net_end) g += l2 * wv[k];
            const uint32_t st = sv[k] + 1;
            if (st != c_st) {
                c = adam_bias(st, beta1, beta2);
                c_st = st;
            }
            adam_apply(g, c, lr, beta1, beta2, eps, wv[k], mv[k], vv[k]);
            sv[k] = st;
        }
        st_nt(w + i4, wv);
        st_nt(m + i4, mv);
        st_nt(v + i4, vv);
        st_nt(steps + i4, sv);
        if (w_lp) LowPrecision4<TL>::store(w_lp, i4, wv);
        *reinterpret_cast<float4*>(grad + i4) = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    } else {
        for (uint32_t i = i4; i < n; ++i) {
            const float g = grad[i];
            adam_one<TL>(i, g, n_grid, n_net_end, inv_loss_scale, lr, beta1, beta2, eps, l2, w, w_lp, m, v, steps);
            if (g != 0.0f) grad[i] = 0.0f;
        }
    }
}

template <typename TL>
__global__ void cast_kernel(uint64_t n, const float* __restrict__ src, TL* __restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (TL)src[i];
}

}  // namespace tn

using namespace tn;

// lp_dtype: 0 = no low-precision copy (plain fp32 training), 1 = fp16 copy.
// Parameters [0, n_grid) are hash-table entries, [n_grid, n_net_end) network
// weights (L2-regularised), the rest (if any) neither.
extern "C" int tn_adam_step(uint32_t n, uint32_t n_grid, uint32_t n_net_end, float loss_scale,
                            const uint32_t* scale_by_rays_info, float learning_rate, const uint32_t* d_step,
                            float lr_decay, float beta1, float beta2, float epsilon, float l2_reg, float* weights,
                            int lp_dtype, void* weights_lp, float* gradients, float* first_moments,
                            float* second_moments, uint32_t* param_steps, hipStream_t stream) {
    TN_CHECK_ARG(weights && gradients && first_moments && second_moments && param_steps && loss_scale > 0.0f);
    TN_CHECK_ARG(n_grid <= n_net_end && n_net_end <= n && (lp_dtype == 0 || weights_lp));
    // 16-byte vector access (8 bytes for the fp16 copy)
    auto aligned = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; };
    TN_CHECK_ARG(aligned(weights, 16) && aligned(gradients, 16) && aligned(first_moments, 16) &&
                 aligned(second_moments, 16) && aligned(param_steps, 16) && (lp_dtype == 0 || aligned(weights_lp, 8)));
    if (n == 0) return TN_OK;
    const unsigned tb = (unsigned)tune_value(TN_TUNE_ADAM_BLOCK);
    const unsigned blocks = blocks_for((n + 3) / 4, tb);
    const float inv = 1.0f / loss_scale;
    if (lp_dtype == 0)
        adam_kernel<float><<<blocks, tb, 0, stream>>>(n, n_grid, n_net_end, inv, scale_by_rays_info, learning_rate,
                                                       d_step, lr_decay, beta1, beta2, epsilon, l2_reg, weights,
                                                       nullptr, gradients, first_moments, second_moments, param_steps);
    else if (lp_dtype == 1)
        adam_kernel<_Float16><<<blocks, tb, 0, stream>>>(n, n_grid, n_net_end, inv, scale_by_rays_info,
                                                          learning_rate, d_step, lr_decay, beta1, beta2, epsilon,
                                                          l2_reg, weights, static_cast<_Float16*>(weights_lp),
                                                          gradients, first_moments, second_moments, param_steps);
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// fp32 -> low-precision copy (reference copy_and_cast, src/utils/parallel-utils.cuh:60-82)
extern "C" int tn_cast_params(uint64_t n, const float* src, int lp_dtype, void* dst, hipStream_t stream) {
    TN_CHECK_ARG(src && dst);
    if (n == 0) return TN_OK;
    if (lp_dtype == 1)
        cast_kernel<_Float16><<<blocks_for(n, 256), 256, 0, stream>>>(n, src, static_cast<_Float16*>(dst));
    else
        return TN_ERR_INVALID_ARGUMENT;
    TN_CHECK_LAUNCH();
    return TN_OK;
}
