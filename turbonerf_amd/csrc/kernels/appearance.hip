// Appearance embedding of the colour MLP as a per-image bias (the struct
// tn_appearance_t documents the formulation). The reference concatenates
// emb[k] to every sample's colour input (src/core/nerf-network.cu:494-520);
// here W3e emb[k] is formed once per image, and the backward pass sums the
// colour layer's pre-activation gradient per image first:
//   S[k]        = sum over samples s of image k of dz[:, s]   (64 values)
//   dL/dW3e     = sum_k S[k] emb[k]^T
//   dL/demb[k]  = W3e^T S[k]
// which equals summing the per-sample products. Samples of one image are
// contiguous in a training batch (rays are assigned to images in order and
// samples follow their rays), so each image's sum is over one range of
// samples, found by binary search; every sum runs in a fixed order, so the
// gradients are the same bit for bit from run to run.
#include "common.hpp"

namespace tn {

// bias[k][r] = sum_d W3e[r][d] emb[k][d], one thread per (k, r)
__global__ void appearance_bias_kernel(uint32_t n, const float* __restrict__ w3e, const float* __restrict__ emb,
                                       float* __restrict__ bias) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 64u) return;
    const uint32_t k = i / 64u, r = i % 64u;
    float s = 0.0f;
#pragma unroll
    for (int d = 0; d < TN_APPEARANCE_DIMS; ++d) s = fmaf(w3e[r * TN_APPEARANCE_DIMS + d], emb[k * TN_APPEARANCE_DIMS + d], s);
    bias[i] = s;
}

// first index in [0, n) with id[i] >= k (ids non-decreasing)
TN_DEVICE uint32_t lower_bound_id(const uint32_t* __restrict__ id, uint32_t n, uint32_t k) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (id[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// S[k] for image k = blockIdx.x: wave w sums rows 16w..16w+15, its lanes
// striding over the image's samples, then a fixed butterfly over the lanes
__global__ void __launch_bounds__(256) appearance_sum_kernel(uint32_t n_max, const uint32_t* __restrict__ info,
                                                             uint32_t B, const uint32_t* __restrict__ id,
                                                             const float* __restrict__ dz, float* __restrict__ S) {
    const uint32_t k = blockIdx.x;
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t lo = lower_bound_id(id, n, k), hi = lower_bound_id(id, n, k + 1);
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    float acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0f;
    for (uint32_t s = lo + lane; s < hi; s += 64) {
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] += dz[(uint64_t)(16 * w + j) * B + s];
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        float v = acc[j];
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if (lane == 0) S[(uint64_t)k * 64 + 16 * w + j] = v;
    }
}

// gradients from S: thread i < 1024 forms dL/dW3e[r][d] (sum over images in
// order); threads beyond form dL/demb[k][d] (sum over the 64 rows in order)
__global__ void appearance_grad_kernel(uint32_t n, const float* __restrict__ S, const float* __restrict__ w3e,
                                       const float* __restrict__ emb, float* __restrict__ grad_w3e,
                                       float* __restrict__ grad_emb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (uint32_t)TN_APPEARANCE_WEIGHTS) {
        const uint32_t r = i / TN_APPEARANCE_DIMS, d = i % TN_APPEARANCE_DIMS;
        float s = 0.0f;
        for (uint32_t k = 0; k < n; ++k) s = fmaf(S[(uint64_t)k * 64 + r], emb[k * TN_APPEARANCE_DIMS + d], s);
        grad_w3e[i] = s;
        return;
    }
    const uint32_t j = i - TN_APPEARANCE_WEIGHTS;
    if (j >= n * TN_APPEARANCE_DIMS) return;
    const uint32_t k = j / TN_APPEARANCE_DIMS, d = j % TN_APPEARANCE_DIMS;
    float s = 0.0f;
#pragma unroll 8
    for (int r = 0; r < 64; ++r) s = fmaf(w3e[r * TN_APPEARANCE_DIMS + d], S[(uint64_t)k * 64 + r], s);
    grad_emb[j] = s;
}

}  // namespace tn

using namespace tn;

extern "C" int tn_appearance_bias(uint32_t n_appearances, const float* w3e, const float* emb, float* bias,
                                  hipStream_t stream) {
    TN_CHECK_ARG(w3e && emb && bias);
    if (n_appearances == 0) return TN_OK;
    appearance_bias_kernel<<<blocks_for((uint64_t)n_appearances * 64, 256), 256, 0, stream>>>(n_appearances, w3e,
                                                                                             emb, bias);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_appearance_workspace_bytes(uint32_t n_appearances) {
    return (size_t)n_appearances * 64 * sizeof(float);
}

extern "C" int tn_appearance_backward(uint32_t n_max, const uint32_t* info, uint32_t stride, const uint32_t* image_id,
                                      const float* dz, uint32_t n_appearances, const float* w3e, const float* emb,
                                      float* grad_w3e, float* grad_emb, void* workspace, size_t workspace_bytes,
                                      hipStream_t stream) {
    TN_CHECK_ARG(image_id && dz && w3e && emb && grad_w3e && grad_emb && workspace && n_max <= stride);
    TN_CHECK_ARG(64ull * stride < (1ull << 32));
    if (n_appearances == 0) return TN_OK;
    if (workspace_bytes < tn_appearance_workspace_bytes(n_appearances)) return TN_ERR_OUT_OF_MEMORY;
    float* S = static_cast<float*>(workspace);
    appearance_sum_kernel<<<n_appearances, 256, 0, stream>>>(n_max, info, stride, image_id, dz, S);
    const uint64_t outs = (uint64_t)TN_APPEARANCE_WEIGHTS + (uint64_t)n_appearances * TN_APPEARANCE_DIMS;
    appearance_grad_kernel<<<blocks_for(outs, 256), 256, 0, stream>>>(n_appearances, S, w3e, emb, grad_w3e,
                                                                      grad_emb);
    TN_CHECK_LAUNCH();
    return TN_OK;
}
