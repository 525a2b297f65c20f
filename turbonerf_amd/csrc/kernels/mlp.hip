// Fully-fused density and colour MLPs on MFMA (reference network shapes at
// src/core/nerf-network.cu:45-121: density 32 -> 64 (ReLU) -> 16, colour
// [16 density outputs | 16 SH-degree-4 coefficients] -> 64 -> 64 -> 3 sigmoid,
// bias-free; forward at :431-545, backward at :668-768, inference at :319-429).
//
// Orientation: a wave processes 32 samples at a time as the N (column)
// dimension of v_mfma_f32_32x32x16_f16 (mixed precision) or
// v_mfma_f32_32x32x2_f32 (plain fp32); features are rows. The fp32
// accumulator tile of one layer is handed to the next layer as its B operand
// without leaving registers: accumulator register r of lane l holds row
// 8*(r/4) + 4*(l/32) + r%4 of column l%32, which is a valid B-operand
// distribution once the next layer's weight columns are permuted the same
// way (dt_in below). Weight gradients dW = delta * act^T use the same MFMA
// with samples as the K dimension after both tiles are transposed through
// LDS; each wave keeps its 12 dW accumulator tiles in registers for all the
// tiles it processes and the workgroup sums them before one partial per
// workgroup goes to memory.
//
// Mixed precision keeps the fp16 operands of the first layer of each MLP
// exact to about 2^-22: the hash features, the hidden layer h1, the density
// outputs and the SH coefficients enter their MFMAs as two fp16 operands,
// hi = fp16(v) and lo = fp16(v - hi), against the same weight operand. Plain
// fp16 rounding of these inputs (2^-11) moves the density and colour enough
// to flip ReLU masks and compositing terms downstream; it was the largest
// part of the gradient error against an FP64 evaluation (about 6e-3 on the
// hash table at 2^18 samples, under 7e-4 with the split). The wider layers
// (h2 -> W4, h3 -> W5) and every backward operand stay single fp16.
//
// Weight layout in the flat parameter buffer (after the hash table):
// W1[64][32] W2[16][64] | W3[64][32] W4[64][64] W5[3][64], row-major [out][in].
#include "common.hpp"

namespace tn {

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef __bf16 b8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

struct PrecF16 {
    using T = _Float16;
    using Op = h8v;
    static constexpr int KC = 16;  // K per MFMA
    static constexpr int SL = 8;   // K values per lane
    static TN_DEVICE f16v mma(Op a, Op b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }
    static TN_DEVICE Op from(const float* v) {
        Op o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (_Float16)v[j];
        return o;
    }
    // v = hi + lo to about 2^-22 relative: hi = fp16(v), lo = fp16(v - hi)
    // (v - hi is exact in fp32)
    static TN_DEVICE void split(const float* v, Op& hi, Op& lo) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const _Float16 a = (_Float16)v[j];
            hi[j] = a;
            lo[j] = (_Float16)(v[j] - (float)a);
        }
    }
};

struct PrecF32 {
    using T = float;
    using Op = float;
    static constexpr int KC = 2;
    static constexpr int SL = 1;
    static TN_DEVICE f16v mma(Op a, Op b, f16v c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
    static TN_DEVICE Op from(const float* v) { return v[0]; }
};

constexpr int kNetParams = TN_DENSITY_MLP_PARAMS + TN_COLOR_MLP_PARAMS;  // 9408
constexpr int kW1 = 0, kW2 = 64 * 32, kW3 = kW2 + 16 * 64, kW4 = kW3 + 64 * 32, kW5 = kW4 + 64 * 64;

// K-steps per 32-row accumulator tile
template <class P>
constexpr int kKsPerTile = 32 / P::KC;

// input row fed by K-slot (ks, h, j) when the K dimension comes from
// accumulator tiles (tile-major)
template <class P>
TN_DEVICE int dt_in(int ks, int h, int j) {
    if constexpr (P::SL == 8) {
        const int t = ks >> 1, sub = ks & 1;
        return 32 * t + 16 * sub + (j < 4 ? 4 * h + j : 8 + 4 * h + (j - 4));
    } else {
        const int t = ks >> 4, r = ks & 15;
        return 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
    }
}

// input row fed by K-slot (ks, h, j) in natural order
template <class P>
TN_DEVICE int nat_in(int ks, int h, int j) {
    return ks * P::KC + h * P::SL + j;
}

// B operand for K-step ks of an accumulator tile
template <class P>
TN_DEVICE typename P::Op dt_op(const f16v& D, int ks) {
    if constexpr (P::SL == 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = D[8 * ks + j];
        return P::from(v);
    } else {
        return D[ks];
    }
}

TN_DEVICE f16v relu(f16v v) {
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = fmaxf(v[r], 0.0f);
    return v;
}

TN_DEVICE f16v zero16() {
    f16v z;
#pragma unroll
    for (int r = 0; r < 16; ++r) z[r] = 0.0f;
    return z;
}

// ---------------------------------------------------------------------------
// Weight staging. A layer's A operand is described by: source matrix offset,
// its (n_out, n_in), whether the A matrix is W or W^T, the number of 32-row M
// tiles and K-steps, and how K-steps map to input rows (kind).
// kind 0: natural; 1: accumulator tiles; 2: colour input (K-steps of the
// first 16 rows come from the density-output tile, the rest natural).
// ---------------------------------------------------------------------------
struct LayerDesc {
    int woff, n_out, n_in, transpose, n_mt, n_ks, kind;
};

template <class P>
TN_DEVICE int k_index(const LayerDesc& L, int ks, int h, int j) {
    if (L.kind == 0) return nat_in<P>(ks, h, j);
    if (L.kind == 1) return dt_in<P>(ks, h, j);
    // kind 2: half the K-steps from the density-output tile (rows 0..15)
    const int half = kKsPerTile<P> / 2;
    if (ks < half) return dt_in<P>(ks, h, j);
    return 16 + nat_in<P>(ks - half, h, j);
}

// pre-arranged A operands: [mt][ks][lane][SL]
template <class P>
TN_DEVICE int stage_layer(typename P::T* dst, const typename P::T* __restrict__ params, const LayerDesc& L) {
    const int total = L.n_mt * L.n_ks * 64 * P::SL;
    for (int e = threadIdx.x; e < total; e += blockDim.x) {
        const int j = e % P::SL;
        const int lane = (e / P::SL) % 64;
        const int ks = (e / (P::SL * 64)) % L.n_ks;
        const int mt = e / (P::SL * 64 * L.n_ks);
        const int m = 32 * mt + (lane & 31);
        const int k = k_index<P>(L, ks, lane >> 5, j);
        const int row = L.transpose ? k : m;
        const int col = L.transpose ? m : k;
        typename P::T v = (typename P::T)0.0f;
        if (row < L.n_out && col < L.n_in) v = params[L.woff + row * L.n_in + col];
        dst[e] = v;
    }
    return total;
}

template <class P>
TN_DEVICE typename P::Op load_a(const typename P::T* lds, const LayerDesc& L, int mt, int ks, int lane) {
    const typename P::T* p = lds + ((mt * L.n_ks + ks) * 64 + lane) * P::SL;
    return *reinterpret_cast<const typename P::Op*>(p);
}

template <class P>
struct Net {
    static constexpr int KT = kKsPerTile<P>;
    // forward layers
    static constexpr LayerDesc L1 = {kW1, 64, 32, 0, 2, 32 / P::KC, 0};
    static constexpr LayerDesc L2 = {kW2, 16, 64, 0, 1, 64 / P::KC, 1};
    static constexpr LayerDesc L3 = {kW3, 64, 32, 0, 2, 32 / P::KC, 2};
    static constexpr LayerDesc L4 = {kW4, 64, 64, 0, 2, 64 / P::KC, 1};
    static constexpr LayerDesc L5 = {kW5, 3, 64, 0, 1, 64 / P::KC, 1};
    // backward (transposed) layers
    static constexpr LayerDesc B5 = {kW5, 3, 64, 1, 2, (P::SL == 8 ? 1 : 4), 1};
    static constexpr LayerDesc B4 = {kW4, 64, 64, 1, 2, 64 / P::KC, 1};
    static constexpr LayerDesc B3 = {kW3, 64, 32, 1, 1, 64 / P::KC, 1};
    static constexpr LayerDesc B2 = {kW2, 16, 64, 1, 2, (P::SL == 8 ? 1 : 8), 1};
    static constexpr LayerDesc B1 = {kW1, 64, 32, 1, 1, 64 / P::KC, 1};

    static constexpr int size(const LayerDesc& L) { return L.n_mt * L.n_ks * 64 * P::SL; }
    static constexpr int kFwdElems = size(L1) + size(L2) + size(L3) + size(L4) + size(L5);
    static constexpr int kBwdElems = size(B5) + size(B4) + size(B3) + size(B2) + size(B1);
};

template <class P>
struct FwdLds {
    typename P::T *l1, *l2, *l3, *l4, *l5;
};

template <class P>
TN_DEVICE FwdLds<P> stage_forward(typename P::T* lds, const typename P::T* __restrict__ params) {
    using N = Net<P>;
    FwdLds<P> f;
    typename P::T* p = lds;
    f.l1 = p; p += stage_layer<P>(p, params, N::L1);
    f.l2 = p; p += stage_layer<P>(p, params, N::L2);
    f.l3 = p; p += stage_layer<P>(p, params, N::L3);
    f.l4 = p; p += stage_layer<P>(p, params, N::L4);
    f.l5 = p; p += stage_layer<P>(p, params, N::L5);
    return f;
}

// y = A . x over K-steps from accumulator tiles x[]
template <class P, int NT_IN>
TN_DEVICE f16v gemm_dt(const typename P::T* lds, const LayerDesc& L, int mt, const f16v (&x)[NT_IN], int lane,
                       int n_ks) {
    f16v acc = zero16();
#pragma unroll
    for (int ks = 0; ks < n_ks; ++ks) {
        const typename P::Op a = load_a<P>(lds, L, mt, ks, lane);
        acc = P::mma(a, dt_op<P>(x[ks / kKsPerTile<P>], ks % kKsPerTile<P>), acc);
    }
    return acc;
}

// An activation tile kept as the next layer's B operands (network precision,
// two fp16 values per register in mixed precision): K-step ks holds
// accumulator registers [SL*ks, SL*ks + SL) of the fp32 tile.
template <class P>
struct OpTile {
    typename P::Op k[kKsPerTile<P>];
};

template <class P>
TN_DEVICE OpTile<P> to_ops(const f16v& D) {
    OpTile<P> t;
#pragma unroll
    for (int ks = 0; ks < kKsPerTile<P>; ++ks) t.k[ks] = dt_op<P>(D, ks);
    return t;
}

// to_ops(relu(D)). Mixed precision rounds first and takes the maximum with
// zero on the fp16 pairs (a packed max per two values instead of a max per
// value and its NaN-quieting twin); fp16 rounding is monotonic and keeps the
// sign, so the values are those of rounding relu(D), up to the sign of
// zero, which no later use distinguishes (relu' tests the magnitude bits,
// and an MFMA operand of -0 adds nothing, as +0 does)
template <class P>
TN_DEVICE OpTile<P> relu_ops(const f16v& D) {
    if constexpr (P::SL == 8) {
        OpTile<P> t = to_ops<P>(D);
        const typename P::Op z = {};
#pragma unroll
        for (int ks = 0; ks < kKsPerTile<P>; ++ks) t.k[ks] = __builtin_elementwise_max(t.k[ks], z);
        return t;
    } else {
        return to_ops<P>(relu(D));
    }
}

// lo part of a tile whose hi part is `hi` (mixed precision only)
template <class P>
TN_DEVICE OpTile<P> lo_ops(const f16v& D, const OpTile<P>& hi) {
    OpTile<P> t;
    if constexpr (P::SL == 8) {
#pragma unroll
        for (int ks = 0; ks < kKsPerTile<P>; ++ks)
#pragma unroll
            for (int j = 0; j < 8; ++j) t.k[ks][j] = (_Float16)(D[8 * ks + j] - (float)hi.k[ks][j]);
    }
    return t;
}

// accumulator register r of a kept tile
template <class P>
TN_DEVICE float op_get(const OpTile<P>& t, int r) {
    if constexpr (P::SL == 8) return (float)t.k[r / 8][r % 8];
    else return t.k[r];
}

// y = A . x over K-steps from kept tiles x[]
template <class P, int NT_IN>
TN_DEVICE f16v gemm_op(const typename P::T* lds, const LayerDesc& L, int mt, const OpTile<P> (&x)[NT_IN], int lane,
                       int n_ks) {
    f16v acc = zero16();
#pragma unroll
    for (int ks = 0; ks < n_ks; ++ks)
        acc = P::mma(load_a<P>(lds, L, mt, ks, lane), x[ks / kKsPerTile<P>].k[ks % kKsPerTile<P>], acc);
    return acc;
}

// y = A . (x + x_lo) over K-steps from kept tiles (mixed precision)
template <class P, int NT_IN>
TN_DEVICE f16v gemm_op2(const typename P::T* lds, const LayerDesc& L, int mt, const OpTile<P> (&x)[NT_IN],
                        const OpTile<P> (&xlo)[NT_IN], int lane, int n_ks) {
    f16v acc = zero16();
#pragma unroll
    for (int ks = 0; ks < n_ks; ++ks) {
        const typename P::Op a = load_a<P>(lds, L, mt, ks, lane);
        acc = P::mma(a, x[ks / kKsPerTile<P>].k[ks % kKsPerTile<P>], acc);
        acc = P::mma(a, xlo[ks / kKsPerTile<P>].k[ks % kKsPerTile<P>], acc);
    }
    return acc;
}

// Real SH basis to degree 4 (16 coefficients) of direction d = 2*in - 1
// (Sloan, "Efficient Spherical Harmonic Evaluation", JCGT 2013).
TN_DEVICE void sh4(float ix, float iy, float iz, float* o) {
    const float x = 2.0f * ix - 1.0f, y = 2.0f * iy - 1.0f, z = 2.0f * iz - 1.0f;
    const float xx = x * x, yy = y * y, zz = z * z;
    o[0] = 0.28209479177387814f;
    o[1] = -0.48860251190291987f * y;
    o[2] = 0.48860251190291987f * z;
    o[3] = -0.48860251190291987f * x;
    o[4] = 1.0925484305920792f * x * y;
    o[5] = -1.0925484305920792f * y * z;
    o[6] = 0.94617469575755997f * zz - 0.31539156525251999f;
    o[7] = -1.0925484305920792f * x * z;
    o[8] = 0.54627421529603959f * (xx - yy);
    o[9] = 0.59004358992664352f * y * (-3.0f * xx + yy);
    o[10] = 2.8906114426405538f * x * y * z;
    o[11] = 0.45704579946446572f * y * (1.0f - 5.0f * zz);
    o[12] = 0.3731763325901154f * z * (5.0f * zz - 3.0f);
    o[13] = 0.45704579946446572f * x * (1.0f - 5.0f * zz);
    o[14] = 1.4453057213202769f * z * (xx - yy);
    o[15] = 0.59004358992664352f * x * (-xx + 3.0f * yy);
}

// this lane's appearance row: the bias row of its sample's image (row 0 for
// every sample when no image ids are given), or null without an embedding
TN_DEVICE const float* app_row(const tn_appearance_t& app, bool valid, uint32_t sample) {
    if (!app.bias) return nullptr;
    const uint32_t k = app.image_id && valid ? app.image_id[sample] : 0u;
    return app.bias + 64u * (k < app.n_appearances ? k : 0u);
}

// Activations of one 32-sample tile (the forward pass; the backward pass
// recomputes it instead of storing activations).
template <class P>
struct Tile {
    typename P::Op x[32 / P::KC];   // hash features, natural order
    typename P::Op sh[16 / P::KC];  // SH coefficients, natural order
    OpTile<P> h1[2], s, h2[2], h3[2];
    float s0;  // raw density output (row 0 of s)
    f16v o;    // colour logits (rows 0..2)
};

// lo parts of the split inputs (mixed precision; used by the first layers'
// MFMAs only, so they live just through the forward pass)
template <class P>
struct TileLo {
    typename P::Op x[32 / P::KC];
    typename P::Op sh[16 / P::KC];
};

// 32-bit loads and stores through a buffer resource over [base, base +
// bytes): the per-lane byte offset goes in a VGPR and the wave-uniform one
// (a row times the row stride) in an SGPR, so the row loops below spend no
// vector arithmetic on 64-bit addresses. Descriptor word 3 of the gfx9
// family: 32-bit data format, no swizzle, no stride (byte-range checked).
TN_DEVICE __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
TN_DEVICE float buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t uni_off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)lane_off, (int)uni_off, 0));
}
TN_DEVICE void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t uni_off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)lane_off, (int)uni_off, 0);
}

// Inputs of one tile. Lanes past the batch load sample 0 and zero what they
// loaded (no divergent branches around the loads); element offsets are
// 32-bit (the host checks that 32 rows of features span under 2^32 bytes).
// kSplit: the first layers' inputs as fp16 hi + lo pairs (mixed-precision
// training); otherwise rounded once. fp16 features (TF = _Float16, for
// inference) are the operands as loaded.
template <class P, typename TF, bool kSplit = (P::SL == 8)>
TN_DEVICE void load_inputs(Tile<P>& tl, TileLo<P>& lo, bool valid, uint32_t sample, int h,
                           const TF* __restrict__ feat, uint32_t FS, const float* __restrict__ dir, uint32_t B,
                           bool need_sh) {
    const uint32_t sl = valid ? sample : 0u;
    static_assert(sizeof(TF) == 4 || (sizeof(TF) == 2 && P::SL == 8 && !kSplit),
                  "features are fp32, or fp16 unsplit for the fp16 network");
    // row nat_in(ks, h, j) = (ks*KC + j) + h*SL: the first part is uniform
    constexpr uint32_t kFB = sizeof(TF);
    const __amdgpu_buffer_rsrc_t fr = buf_rsrc(feat, 32u * FS * kFB);
    const uint32_t lane_off = ((uint32_t)(h * P::SL) * FS + sl) * kFB;
#pragma unroll
    for (int ks = 0; ks < 32 / P::KC; ++ks) {
        if constexpr (sizeof(TF) == 2) {
            typename P::Op o;
#pragma unroll
            for (int j = 0; j < P::SL; ++j) {
                const uint16_t x = __builtin_amdgcn_raw_buffer_load_b16(fr, (int)lane_off,
                                                                        (int)((uint32_t)(ks * P::KC + j) * FS * 2u), 0);
                o[j] = valid ? __builtin_bit_cast(_Float16, x) : (_Float16)0.0f;
            }
            tl.x[ks] = o;
        } else {
            float v[P::SL];
#pragma unroll
            for (int j = 0; j < P::SL; ++j) {
                const float x = buf_ld(fr, lane_off, (uint32_t)(ks * P::KC + j) * FS * 4u);
                v[j] = valid ? x : 0.0f;
            }
            if constexpr (kSplit) P::split(v, tl.x[ks], lo.x[ks]);
            else tl.x[ks] = P::from(v);
        }
    }
    if (need_sh) {
        float shv[16];
        sh4(dir[sl], dir[sl + B], dir[sl + 2 * B], shv);
        if constexpr (P::SL == 8 && !kSplit) {
            // the lane half's eight coefficients, selected as values
            const typename P::Op o0 = P::from(shv), o1 = P::from(shv + 8);
            const typename P::Op z = {};
            tl.sh[0] = valid ? (h ? o1 : o0) : z;
        } else if constexpr (P::SL == 8) {
            // both halves' operands are formed and the lane half's pair is
            // selected as values: an array index that depends on h becomes
            // a compare-and-select chain over all 16 coefficients
            typename P::Op hi0, lo0, hi1, lo1;
            P::split(shv, hi0, lo0);
            P::split(shv + 8, hi1, lo1);
            const typename P::Op z = {};
            tl.sh[0] = valid ? (h ? hi1 : hi0) : z;
            lo.sh[0] = valid ? (h ? lo1 : lo0) : z;
        } else {
            if (!valid)
                for (int k = 0; k < 16; ++k) shv[k] = 0.0f;
#pragma unroll
            for (int ks = 0; ks < 16 / P::KC; ++ks) {
                float v[P::SL];
#pragma unroll
                for (int j = 0; j < P::SL; ++j) v[j] = shv[nat_in<P>(ks, h, j)];
                tl.sh[ks] = P::from(v);
            }
        }
    }
}

// row of accumulator register r in lane half h of a 32-row tile t
TN_DEVICE int acc_row(int t, int r, int h) { return 32 * t + 8 * (r >> 2) + 4 * h + (r & 3); }

// accumulator tile t holding rows 32t.. of a 64-vector (4 float4 loads)
TN_DEVICE f16v load_rows(const float* __restrict__ v, int t, int h) {
    f16v a;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 f = *reinterpret_cast<const float4*>(v + 32 * t + 8 * q + 4 * h);
        a[4 * q] = f.x;
        a[4 * q + 1] = f.y;
        a[4 * q + 2] = f.z;
        a[4 * q + 3] = f.w;
    }
    return a;
}

// Forward of one tile. app_bias (may be null): this lane's sample's
// appearance row, W3[:, 32:48] times its image's embedding, which starts the
// colour MLP's first-layer accumulator.
template <class P, bool kSplit = (P::SL == 8)>
TN_DEVICE void forward_tile(Tile<P>& tl, const TileLo<P>& lo, const FwdLds<P>& W, int lane, bool color,
                            const float* __restrict__ app_bias = nullptr) {
    using N = Net<P>;
    OpTile<P> h1lo[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        f16v acc = zero16();
#pragma unroll
        for (int ks = 0; ks < N::L1.n_ks; ++ks) {
            const typename P::Op a = load_a<P>(W.l1, N::L1, mt, ks, lane);
            acc = P::mma(a, tl.x[ks], acc);
            if constexpr (kSplit) acc = P::mma(a, lo.x[ks], acc);
        }
        const f16v r = relu(acc);
        tl.h1[mt] = to_ops<P>(r);
        if constexpr (kSplit) h1lo[mt] = lo_ops<P>(r, tl.h1[mt]);
    }
    f16v s;
    if constexpr (kSplit) s = gemm_op2<P, 2>(W.l2, N::L2, 0, tl.h1, h1lo, lane, N::L2.n_ks);
    else s = gemm_op<P, 2>(W.l2, N::L2, 0, tl.h1, lane, N::L2.n_ks);
    tl.s0 = s[0];
    tl.s = to_ops<P>(s);
    if (!color) return;
    const OpTile<P> slo = lo_ops<P>(s, tl.s);
    constexpr int half = kKsPerTile<P> / 2;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        f16v acc = app_bias ? load_rows(app_bias, mt, lane >> 5) : zero16();
#pragma unroll
        for (int ks = 0; ks < half; ++ks) {
            const typename P::Op a = load_a<P>(W.l3, N::L3, mt, ks, lane);
            acc = P::mma(a, tl.s.k[ks], acc);
            if constexpr (kSplit) acc = P::mma(a, slo.k[ks], acc);
        }
#pragma unroll
        for (int ks = 0; ks < 16 / P::KC; ++ks) {
            const typename P::Op a = load_a<P>(W.l3, N::L3, mt, half + ks, lane);
            acc = P::mma(a, tl.sh[ks], acc);
            if constexpr (kSplit) acc = P::mma(a, lo.sh[ks], acc);
        }
        tl.h2[mt] = relu_ops<P>(acc);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) tl.h3[mt] = relu_ops<P>(gemm_op<P, 2>(W.l4, N::L4, mt, tl.h2, lane, N::L4.n_ks));
    tl.o = gemm_op<P, 2>(W.l5, N::L5, 0, tl.h3, lane, N::L5.n_ks);
}

TN_DEVICE float sigmoidf(float x) { return 1.0f / (1.0f + __expf(-x)); }

constexpr int kMlpWaves = 4;

template <class P, typename TF, bool kSplit = (P::SL == 8)>
__global__ void __launch_bounds__(64 * kMlpWaves)
mlp_forward_kernel(uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t FS,
                   const typename P::T* __restrict__ params, const TF* __restrict__ feat,
                   const float* __restrict__ dir, int color, typename P::T* __restrict__ density_out,
                   typename P::T* __restrict__ rgb_out, uint32_t OS, tn_appearance_t app) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    typename P::T* lds = reinterpret_cast<typename P::T*>(smem_raw);
    const FwdLds<P> W = stage_forward<P>(lds, params);
    __syncthreads();
    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t n_tiles = (n + 31) / 32;
    const int lane = (int)__lane_id();
    const int h = lane >> 5;
    const uint32_t wave = blockIdx.x * kMlpWaves + (threadIdx.x >> 6);
    const uint32_t n_waves = gridDim.x * kMlpWaves;
    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const uint32_t sample = t * 32 + (lane & 31);
        const bool valid = sample < n;
        Tile<P> tl;
        TileLo<P> lo;
        load_inputs<P, TF, kSplit>(tl, lo, valid, sample, h, feat, FS, dir, B, color != 0);
        forward_tile<P, kSplit>(tl, lo, W, lane, color != 0, app_row(app, valid, sample));
        if (valid && h == 0) {
            density_out[sample] = (typename P::T)tl.s0;
            if (color) {
                rgb_out[sample] = (typename P::T)sigmoidf(tl.o[0]);
                rgb_out[sample + OS] = (typename P::T)sigmoidf(tl.o[1]);
                rgb_out[sample + 2 * OS] = (typename P::T)sigmoidf(tl.o[2]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// backward
// ---------------------------------------------------------------------------
template <class P>
struct BwdLds {
    typename P::T *b5, *b4, *b3, *b2, *b1;
};

template <class P>
TN_DEVICE BwdLds<P> stage_backward(typename P::T* lds, const typename P::T* __restrict__ params) {
    using N = Net<P>;
    BwdLds<P> b;
    typename P::T* p = lds;
    b.b5 = p; p += stage_layer<P>(p, params, N::B5);
    b.b4 = p; p += stage_layer<P>(p, params, N::B4);
    b.b3 = p; p += stage_layer<P>(p, params, N::B3);
    b.b2 = p; p += stage_layer<P>(p, params, N::B2);
    b.b1 = p; p += stage_layer<P>(p, params, N::B1);
    return b;
}

// Backward. A workgroup of kMlpWaves = 4 waves takes 128 samples at a time,
// wave w the 32-sample tile w. Each wave recomputes its tile's forward pass,
// keeps the activations as operand tiles and runs the delta chain down to
// dfeat. Weight gradients dW = delta * act^T take the samples as K: each
// layer's deltas and activations of all four tiles go to LDS, and the four
// waves share the layer's 32x32 dW tiles -- with four tiles (W4) one each,
// with two tiles (W1, W2, W3, W5) each tile's 128 samples split between two
// waves -- so a wave accumulates five dW tiles across the batch instead of
// twelve.
//
// The MFMA operands of dW hold 8 samples of one row per lane, the
// accumulator tiles 16 rows of one sample, so the group buffers transpose.
// Mixed precision stores a sample's rows contiguously ([sample][row], a
// lane's four consecutive accumulator rows as one 8-byte write) and reads
// the operands with gfx950's transposing LDS load ds_read_b64_tr_b16: in
// each 16-lane group, lane t's element e is the 16-bit value at (address of
// lane 4e + t/4) + t%4, so lanes that address (sample e, rows 4q..4q+3) give
// lane 4q + i of the group the row 4q + i of samples 0..3. Plain fp32 keeps
// [row][sample] buffers written one value per lane.
constexpr int kGrp = 32 * kMlpWaves;  // samples per workgroup step
static_assert(kMlpWaves == 4, "the dW split assumes four waves");

// fp32 buffers: [row][sample], row stride 129 (the +1 spreads banks)
template <class P>
constexpr int kGW = kGrp + 1;
// fp16 buffers: [sample][row], sample stride 100 fp16 = 200 bytes (64 rows
// and padding): 200 = 8 * 25 with 25 odd puts the 8-byte writes of 32
// consecutive samples in 32 distinct bank pairs, and the four samples of a
// transposing read 7 or more 8-byte slots apart
constexpr int kTrRows = 100;

template <class P>
constexpr size_t grp_buf_elems() {
    return P::SL == 8 ? (size_t)kGrp * kTrRows : (size_t)64 * kGW<P>;
}

typedef _Float16 h4v __attribute__((ext_vector_type(4)));
typedef short s4v __attribute__((ext_vector_type(4)));

// this lane's first element of a [sample][row] buffer: sample 32w + (lane &
// 31), rows from 4 * (lane half)
template <class P>
TN_DEVICE typename P::T* tr_put_base(typename P::T* buf, int w, int lane) {
    return buf + (32 * w + (lane & 31)) * kTrRows + 4 * ((lane >> 5) & 1);
}

// tile t of a wave's samples (samples 32w..32w+31); HALF: rows 32t..32t+15
// only (registers 0..7)
template <class P, bool HALF = false>
TN_DEVICE void put_grp(typename P::T* buf, int t, int w, const OpTile<P>& A, int lane) {
    if constexpr (P::SL == 8) {
        // operand k[ks] holds accumulator registers 8ks..8ks+7: rows
        // 32t + 16ks + 4h + (0..3) and 32t + 16ks + 8 + 4h + (0..3)
        typename P::T* p = tr_put_base<P>(buf, w, lane) + 32 * t;
#pragma unroll
        for (int ks = 0; ks < (HALF ? 1 : 2); ++ks) {
            *reinterpret_cast<h4v*>(p + 16 * ks) = __builtin_shufflevector(A.k[ks], A.k[ks], 0, 1, 2, 3);
            *reinterpret_cast<h4v*>(p + 16 * ks + 8) = __builtin_shufflevector(A.k[ks], A.k[ks], 4, 5, 6, 7);
        }
    } else {
        const int n = 32 * w + (lane & 31), h = (lane >> 5) & 1;
#pragma unroll
        for (int r = 0; r < (HALF ? 8 : 16); ++r) buf[acc_row(t, r, h) * kGW<P> + n] = op_get<P>(A, r);
    }
}

template <class P>
TN_DEVICE void put_grp(typename P::T* buf, int t, int w, const f16v& D, int lane) {
    if constexpr (P::SL == 8) {
        put_grp<P>(buf, t, w, to_ops<P>(D), lane);
    } else {
        const int n = 32 * w + (lane & 31), h = (lane >> 5) & 1;
#pragma unroll
        for (int r = 0; r < 16; ++r) buf[acc_row(t, r, h) * kGW<P> + n] = D[r];
    }
}

// natural-order operands x[ks] as rows base + ks*KC + h*SL + j
template <class P, int NKS>
TN_DEVICE void put_grp_nat(typename P::T* buf, int base, int w, const typename P::Op (&x)[NKS], int lane) {
    if constexpr (P::SL == 8) {
        // rows base + 16ks + 8h + (0..7): two 8-byte writes per K-step (a
        // sample's rows start 8-byte aligned only)
        typename P::T* p = tr_put_base<P>(buf, w, lane) + base + 4 * ((lane >> 5) & 1);
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
            *reinterpret_cast<h4v*>(p + 16 * ks) = __builtin_shufflevector(x[ks], x[ks], 0, 1, 2, 3);
            *reinterpret_cast<h4v*>(p + 16 * ks + 4) = __builtin_shufflevector(x[ks], x[ks], 4, 5, 6, 7);
        }
    } else {
        const int n = 32 * w + (lane & 31), h = (lane >> 5) & 1;
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) buf[(base + nat_in<P>(ks, h, 0)) * kGW<P> + n] = x[ks];
    }
}

// operand holding group samples [ks*KC + h*SL, +SL) of row 32*t + (lane & 31)
template <class P>
TN_DEVICE typename P::Op grp_rows(const typename P::T* buf, int t, int ks, int lane) {
    if constexpr (P::SL == 8) {
        // group G = lane / 16 reads rows 32t + 16(G % 2) + (0..15) of samples
        // 16ks + 8(G / 2) + (0..7); its lane 4e + q addresses sample e (and
        // e + 4 in the second read), rows 4q..4q+3 of those 16
        const int G = lane >> 4, jj = lane & 15;
        const typename P::T* p =
            buf + (16 * ks + 8 * (G >> 1) + (jj >> 2)) * kTrRows + 32 * t + 16 * (G & 1) + 4 * (jj & 3);
        typedef s4v __attribute__((address_space(3))) * lds_s4v;
        const s4v a = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4v)(p));
        const s4v b = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_s4v)(p + 4 * kTrRows));
        return __builtin_bit_cast(typename P::Op, __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7));
    } else {
        const int row = 32 * t + (lane & 31), h = lane >> 5;
        return buf[row * kGW<P> + ks * P::KC + h * P::SL];
    }
}

// C += delta rows (tile mt) x act rows (tile nt)^T over group K-steps [ks0, ks1)
template <class P>
TN_DEVICE f16v dw_tile(f16v c, const typename P::T* dbuf, const typename P::T* abuf, int mt, int nt, int ks0,
                       int ks1, int lane) {
#pragma unroll 4
    for (int ks = ks0; ks < ks1; ++ks) c = P::mma(grp_rows<P>(dbuf, mt, ks, lane), grp_rows<P>(abuf, nt, ks, lane), c);
    return c;
}

template <class P>
TN_DEVICE f16v relu_mask(f16v d, const OpTile<P>& act) {
#pragma unroll
    for (int r = 0; r < 16; ++r) d[r] = op_get<P>(act, r) > 0.0f ? d[r] : 0.0f;
    return d;
}

// to_ops(relu_mask(d, act)) with the mask applied to the packed operands:
// in mixed precision an fp16 activation is > 0 exactly when its magnitude
// bits m are non-zero, and m + 0x7FFF (no carry out of the half, m <= 0x7FFF)
// has its top bit set exactly then; an arithmetic shift by 15 spreads that
// bit over the half, giving an AND mask for d's half (four ops per pair of
// values instead of a convert, a compare and a select per value)
typedef short i16x2 __attribute__((ext_vector_type(2)));
template <class P>
TN_DEVICE OpTile<P> masked_ops(const f16v& d, const OpTile<P>& act) {
    if constexpr (P::SL == 8) {
        OpTile<P> t = to_ops<P>(d);
#pragma unroll
        for (int ks = 0; ks < kKsPerTile<P>; ++ks) {
            typedef uint32_t u4v __attribute__((ext_vector_type(4)));
            u4v dv = __builtin_bit_cast(u4v, t.k[ks]);
            const u4v av = __builtin_bit_cast(u4v, act.k[ks]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const i16x2 top = __builtin_bit_cast(i16x2, (av[q] & 0x7FFF7FFFu) + 0x7FFF7FFFu);
                dv[q] &= __builtin_bit_cast(uint32_t, top >> (i16x2){15, 15});
            }
            t.k[ks] = __builtin_bit_cast(typename P::Op, dv);
        }
        return t;
    } else {
        return to_ops<P>(relu_mask<P>(d, act));
    }
}

// add tile C (rows 32*mt.., cols 32*nt..) of a [n_out][n_in] gradient into LDS
TN_DEVICE void add_dw_tile(float* acc, const f16v& C, int mt, int nt, int n_out, int n_in, int lane) {
    const int col = 32 * nt + (lane & 31), h = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = acc_row(mt, r, h);
        if (row < n_out && col < n_in) acc[row * n_in + col] += C[r];
    }
}

// sets of group buffers: two in mixed precision; plain fp32 (twice the
// bytes) fits one in LDS and waits at a second barrier per layer instead
template <class P>
constexpr int kGrpSets = P::SL == 8 ? 2 : 1;

// after a layer's dW reads: the other set for the next layer, or (one set)
// a barrier before it is written again
template <class P>
TN_DEVICE void next_layer(typename P::T*& d, typename P::T*& a, typename P::T*& dn, typename P::T*& an) {
    if constexpr (kGrpSets<P> == 2) {
        typename P::T* t = d;
        d = dn;
        dn = t;
        t = a;
        a = an;
        an = t;
    } else {
        __syncthreads();
    }
}

template <class P, typename TF, typename TG>
__global__ void __launch_bounds__(64 * kMlpWaves)
mlp_backward_kernel(uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t FS,
                    const typename P::T* __restrict__ params, const TF* __restrict__ feat,
                    const float* __restrict__ dir, const float* __restrict__ sample_t,
                    const TG* __restrict__ dL_drgb, uint32_t GS, const float* __restrict__ dL_ds,
                    const float* __restrict__ dL_ds_extra, float loss_scale, int scale_by_rays,
                    TF* __restrict__ dfeat, float* __restrict__ dw_partial, tn_appearance_t app) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    using T = typename P::T;
    using N = Net<P>;
    constexpr int KS = kGrp / P::KC;  // K-steps of a group
    T* lds = reinterpret_cast<T*>(smem_raw);
    const FwdLds<P> W = stage_forward<P>(lds, params);
    const BwdLds<P> WB = stage_backward<P>(lds + N::kFwdElems, params);
    // two sets of group buffers (deltas and activations, up to 64 rows x 128
    // samples each), used by the layers in turn: a layer's writes go to the
    // set the layer before last read, and every wave has passed the barrier
    // of the layer in between only after its reads of that set, so one
    // barrier per layer orders both the writes before the reads and the
    // reads before the next writes
    T* const bufs = lds + N::kFwdElems + N::kBwdElems;
    constexpr size_t E = grp_buf_elems<P>();
    T* dbuf = bufs;
    T* abuf = bufs + E;
    T* dnext = bufs + (kGrpSets<P> == 2 ? 2 : 0) * E;
    T* anext = bufs + (kGrpSets<P> == 2 ? 3 : 1) * E;
    __syncthreads();

    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t n_grp = (n + kGrp - 1) / kGrp;
    const float ls = scale_by_rays ? loss_scale * ray_pow2(info[0]) : loss_scale;
    const int lane = (int)__lane_id();
    const int h = lane >> 5;
    const int wid = threadIdx.x >> 6;
    // this wave's share of the dW tiles: a four-tile layer's tile (wid >> 1,
    // wid & 1); a two-tile layer's tile wid & 1 over K half wid >> 1
    const int t2 = wid & 1, k0 = (wid >> 1) * (KS / 2), k1 = k0 + KS / 2;
    f16v dW1 = zero16(), dW2 = zero16(), dW3 = zero16(), dW4 = zero16(), dW5 = zero16();

    for (uint32_t grp = blockIdx.x; grp < n_grp; grp += gridDim.x) {
        const uint32_t sample = grp * kGrp + 32 * wid + (lane & 31);
        const bool valid = sample < n;
        const uint32_t sl = valid ? sample : 0u;
        Tile<P> tl;
        TileLo<P> lo;
        load_inputs<P>(tl, lo, valid, sample, h, feat, FS, dir, B, true);
        // per-sample gradients from the compositing, loaded up front
        float g_rgb[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) g_rgb[c] = (float)dL_drgb[sl + c * GS];
        const float g_s = ls * dL_ds[sl];
        const float g_x = dL_ds_extra ? ls * dL_ds_extra[sl] : 0.0f;
        const float t_s = sample_t[sl];
        forward_tile<P>(tl, lo, W, lane, true, app_row(app, valid, sample));

        // dL/do through the sigmoid (rows 0..2 live in lanes 0..31, registers 0..2)
        f16v dO = zero16();
        if (valid && h == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = sigmoidf(tl.o[c]);
                dO[c] = g_rgb[c] * y * (1.0f - y);
            }
        }
        // dH3 = relu'(h3) (W5^T dO); dW5 += dO h3^T
        const f16v dOt[1] = {dO};
        OpTile<P> dH3[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
            dH3[mt] = masked_ops<P>(gemm_dt<P, 1>(WB.b5, N::B5, mt, dOt, lane, N::B5.n_ks), tl.h3[mt]);
        put_grp<P>(dbuf, 0, wid, dO, lane);
        put_grp<P>(abuf, 0, wid, tl.h3[0], lane);
        put_grp<P>(abuf, 1, wid, tl.h3[1], lane);
        __syncthreads();
        dW5 = dw_tile<P>(dW5, dbuf, abuf, 0, t2, k0, k1, lane);
        next_layer<P>(dbuf, abuf, dnext, anext);
        // dH2 = relu'(h2) (W4^T dH3); dW4 += dH3 h2^T
        OpTile<P> dH2[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            const f16v g = gemm_op<P, 2>(WB.b4, N::B4, mt, dH3, lane, N::B4.n_ks);
            dH2[mt] = masked_ops<P>(g, tl.h2[mt]);
            // the appearance bias's gradient per sample (summed per image later)
            if (app.dz && valid) {
                const f16v d = relu_mask<P>(g, tl.h2[mt]);
                const __amdgpu_buffer_rsrc_t zr = buf_rsrc(app.dz, 64u * B * 4u);
                const uint32_t lane_off = ((uint32_t)(4 * h) * B + sample) * 4u;
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    buf_st(zr, lane_off, (uint32_t)(32 * mt + 8 * (r >> 2) + (r & 3)) * B * 4u, d[r]);
            }
        }
        put_grp<P>(dbuf, 0, wid, dH3[0], lane);
        put_grp<P>(dbuf, 1, wid, dH3[1], lane);
        put_grp<P>(abuf, 0, wid, tl.h2[0], lane);
        put_grp<P>(abuf, 1, wid, tl.h2[1], lane);
        __syncthreads();
        dW4 = dw_tile<P>(dW4, dbuf, abuf, wid >> 1, t2, 0, KS, lane);
        next_layer<P>(dbuf, abuf, dnext, anext);
        // dXc = W3^T dH2 (rows 0..15 = d density outputs); dW3 += dH2 xc^T,
        // xc = [density outputs (rows 0..15) | SH (rows 16..31)]
        const f16v dXc = gemm_op<P, 2>(WB.b3, N::B3, 0, dH2, lane, N::B3.n_ks);
        put_grp<P>(dbuf, 0, wid, dH2[0], lane);
        put_grp<P>(dbuf, 1, wid, dH2[1], lane);
        put_grp<P, true>(abuf, 0, wid, tl.s, lane);
        put_grp_nat<P, 16 / P::KC>(abuf, 16, wid, tl.sh, lane);
        __syncthreads();
        dW3 = dw_tile<P>(dW3, dbuf, abuf, t2, 0, k0, k1, lane);
        next_layer<P>(dbuf, abuf, dnext, anext);
        // density-output gradient: row 0 also receives the loss-scaled
        // compositing (and distortion) gradient, then the whole row is scaled
        // by t^2 where t < 1 (reference nerf-network.cu:734-758)
        f16v dS = zero16();
#pragma unroll
        for (int r = 0; r < 8; ++r) dS[r] = dXc[r];  // rows < 16: registers 0..7
        if (valid && h == 0) {
            float g0 = dS[0] + g_s;
            if (dL_ds_extra) g0 += g_x;
            if (t_s < 1.0f) g0 *= t_s * t_s;
            dS[0] = g0;
        }
        // dH1 = relu'(h1) (W2^T dS); dW2 += dS h1^T
        const f16v dSt[1] = {dS};
        OpTile<P> dH1[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
            dH1[mt] = masked_ops<P>(gemm_dt<P, 1>(WB.b2, N::B2, mt, dSt, lane, N::B2.n_ks), tl.h1[mt]);
        put_grp<P>(dbuf, 0, wid, dS, lane);
        put_grp<P>(abuf, 0, wid, tl.h1[0], lane);
        put_grp<P>(abuf, 1, wid, tl.h1[1], lane);
        __syncthreads();
        dW2 = dw_tile<P>(dW2, dbuf, abuf, 0, t2, k0, k1, lane);
        next_layer<P>(dbuf, abuf, dnext, anext);
        // dX = W1^T dH1 -> feature gradients; dW1 += dH1 x^T
        const f16v dX = gemm_op<P, 2>(WB.b1, N::B1, 0, dH1, lane, N::B1.n_ks);
        if (valid) {
            // row acc_row(0, r, h) = (8(r/4) + r%4) + 4h: the first part is uniform
            const __amdgpu_buffer_rsrc_t dr = buf_rsrc(dfeat, 32u * FS * 4u);
            const uint32_t lane_off = ((uint32_t)(4 * h) * FS + sample) * 4u;
#pragma unroll
            for (int r = 0; r < 16; ++r) buf_st(dr, lane_off, (uint32_t)(8 * (r >> 2) + (r & 3)) * FS * 4u, dX[r]);
        }
        put_grp<P>(dbuf, 0, wid, dH1[0], lane);
        put_grp<P>(dbuf, 1, wid, dH1[1], lane);
        put_grp_nat<P, 32 / P::KC>(abuf, 0, wid, tl.x, lane);
        __syncthreads();
        dW1 = dw_tile<P>(dW1, dbuf, abuf, t2, 0, k0, k1, lane);
        next_layer<P>(dbuf, abuf, dnext, anext);
    }

    // workgroup sum of the weight gradients (waves in turn, so the sum order
    // is fixed), then one partial per workgroup
    float* red = reinterpret_cast<float*>(smem_raw);
    for (int e = threadIdx.x; e < kNetParams; e += blockDim.x) red[e] = 0.0f;
    __syncthreads();
    for (int w = 0; w < kMlpWaves; ++w) {
        if (wid == w) {
            add_dw_tile(red + kW5, dW5, 0, t2, 3, 64, lane);
            add_dw_tile(red + kW4, dW4, wid >> 1, t2, 64, 64, lane);
            add_dw_tile(red + kW3, dW3, t2, 0, 64, 32, lane);
            add_dw_tile(red + kW2, dW2, 0, t2, 16, 64, lane);
            add_dw_tile(red + kW1, dW1, t2, 0, 64, 32, lane);
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < kNetParams; e += blockDim.x)
        dw_partial[(uint64_t)blockIdx.x * kNetParams + e] = red[e];
}

// grad[e] (+)= sum over partials in a fixed order: a workgroup owns 64
// consecutive elements, wave w sums partials w, w + 16, ... (one 256-byte row
// segment per partial), then wave 0 adds the 16 wave sums in order.
constexpr int kReduceWaves = 16;

__global__ void __launch_bounds__(64 * kReduceWaves)
reduce_partials_kernel(uint32_t n_part, const float* __restrict__ partial, float* __restrict__ grad, int accumulate) {
    __shared__ float red[kReduceWaves][64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t e = blockIdx.x * 64u + lane;
    float s = 0.0f;
    if (e < (uint32_t)kNetParams) {
        // eight rows in flight, added in row order
        uint32_t p = w;
        for (; p + 7 * kReduceWaves < n_part; p += 8 * kReduceWaves) {
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partial[(uint64_t)(p + u * kReduceWaves) * kNetParams + e];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; p < n_part; p += kReduceWaves) s += partial[(uint64_t)p * kNetParams + e];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && e < (uint32_t)kNetParams) {
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < kReduceWaves; ++k) t += red[k][lane];
        grad[e] = accumulate ? grad[e] + t : t;
    }
}

template <class P>
constexpr size_t fwd_lds_bytes() {
    return (size_t)Net<P>::kFwdElems * sizeof(typename P::T);
}

template <class P>
constexpr size_t bwd_lds_bytes() {
    const size_t w = (size_t)(Net<P>::kFwdElems + Net<P>::kBwdElems) * sizeof(typename P::T);
    const size_t t = 2 * kGrpSets<P> * grp_buf_elems<P>() * sizeof(typename P::T);
    const size_t r = (size_t)kNetParams * sizeof(float);
    return (w + t) > r ? (w + t) : r;
}


static_assert(bwd_lds_bytes<PrecF16>() <= 160 * 1024 && bwd_lds_bytes<PrecF32>() <= 160 * 1024,
              "the backward kernel's LDS must fit one CU (160 KiB)");

// Persistent launches: as many workgroups as fit on the device at once (the
// backward kernel holds one workgroup per CU, the forward two), each looping
// over 32-sample tiles. Fewer workgroups also means fewer weight stagings and
// fewer gradient partials to reduce.
inline unsigned mlp_blocks(uint32_t n_max, unsigned per_cu) {
    const uint32_t tiles = (n_max + 31) / 32;
    const unsigned b = blocks_for(tiles, kMlpWaves);
    const unsigned cap = per_cu * compute_units();
    return b > cap ? cap : (b == 0 ? 1 : b);
}
// the backward pass's partial-sum workspace is sized for the largest
// TN_TUNE_MLP_BWD_BLOCKS_PER_CU
constexpr unsigned kBwdBlocksPerCUMax = 2;

}  // namespace tn

using namespace tn;

// precision: 0 = fp32 (MFMA fp32), 1 = fp16 (MFMA f16, fp32 accumulate).
// params: the network's parameter buffer in that precision (hash table first,
// `net_offset` = its length). features [32][feat_stride] fp32 (both
// precisions), dir [3][stride] fp32 (0.5*d + 0.5). Outputs: density (raw s)
// [stride] and rgb [3][out_stride] in the network precision.
extern "C" int tn_mlp_forward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
                              uint32_t feat_stride, uint32_t out_stride, const void* params, uint32_t net_offset,
                              const void* features, const float* dir, int color, void* density_out, void* rgb_out,
                              const tn_appearance_t* appearance, hipStream_t stream) {
    TN_CHECK_ARG(params && features && density_out && (!color || (dir && rgb_out)));
    const tn_appearance_t app = appearance && color ? *appearance : tn_appearance_t{};
    TN_CHECK_ARG(!app.bias || (app.n_appearances > 0 && (reinterpret_cast<uintptr_t>(app.bias) & 15) == 0));
    TN_CHECK_ARG(n_max <= stride && n_max <= feat_stride && n_max <= out_stride);
    TN_CHECK_ARG(128ull * feat_stride < (1ull << 32) && 3ull * stride < (1ull << 32) && 3ull * out_stride < (1ull << 32));
    if (n_max == 0) return TN_OK;
    const unsigned blocks = mlp_blocks(n_max, (unsigned)tune_value(TN_TUNE_MLP_FWD_BLOCKS_PER_CU));
    if (precision == 1) {
        using P = PrecF16;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        mlp_forward_kernel<P, float><<<blocks, 64 * kMlpWaves, fwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const float*>(features), dir, color,
            static_cast<P::T*>(density_out), static_cast<P::T*>(rgb_out), out_stride, app);
    } else if (precision == 0) {
        using P = PrecF32;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        mlp_forward_kernel<P, P::T><<<blocks, 64 * kMlpWaves, fwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, color,
            static_cast<P::T*>(density_out), static_cast<P::T*>(rgb_out), out_stride, app);
    } else {
        return TN_ERR_INVALID_ARGUMENT;
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// tn_mlp_forward for inference: feat_dtype 1 takes fp16 features (precision
// 1 only) as the first layer's operands, and every layer's inputs are
// rounded to fp16 once, as the reference's fp16 network computes; feat_dtype
// 0 is tn_mlp_forward.
extern "C" int tn_mlp_inference(int precision, int feat_dtype, uint32_t n_max, const uint32_t* info,
                                uint32_t stride, uint32_t feat_stride, uint32_t out_stride, const void* params,
                                uint32_t net_offset, const void* features, const float* dir, int color,
                                void* density_out, void* rgb_out, const tn_appearance_t* appearance,
                                hipStream_t stream) {
    if (feat_dtype == 0)
        return tn_mlp_forward(precision, n_max, info, stride, feat_stride, out_stride, params, net_offset, features,
                              dir, color, density_out, rgb_out, appearance, stream);
    TN_CHECK_ARG(feat_dtype == 1 && precision == 1);
    TN_CHECK_ARG(params && features && density_out && (!color || (dir && rgb_out)));
    const tn_appearance_t app = appearance && color ? *appearance : tn_appearance_t{};
    TN_CHECK_ARG(!app.bias || (app.n_appearances > 0 && (reinterpret_cast<uintptr_t>(app.bias) & 15) == 0));
    TN_CHECK_ARG(n_max <= stride && n_max <= feat_stride && n_max <= out_stride);
    TN_CHECK_ARG(64ull * feat_stride < (1ull << 32) && 3ull * stride < (1ull << 32) && 3ull * out_stride < (1ull << 32));
    if (n_max == 0) return TN_OK;
    const unsigned blocks = mlp_blocks(n_max, (unsigned)tune_value(TN_TUNE_MLP_FWD_BLOCKS_PER_CU));
    using P = PrecF16;
    const auto* w = static_cast<const P::T*>(params) + net_offset;
    mlp_forward_kernel<P, _Float16, false><<<blocks, 64 * kMlpWaves, fwd_lds_bytes<P>(), stream>>>(
        n_max, info, stride, feat_stride, w, static_cast<const _Float16*>(features), dir, color,
        static_cast<P::T*>(density_out), static_cast<P::T*>(rgb_out), out_stride, app);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_mlp_backward_workspace_bytes(uint32_t n_max) {
    return (size_t)mlp_blocks(n_max, kBwdBlocksPerCUMax) * kNetParams * sizeof(float);
}

// Backward of both MLPs for one batch. dL_drgb [3][grad_stride] (loss-scaled,
// in the network precision), dL_ds [stride] fp32 compositing gradient,
// dL_ds_extra (may be null) e.g. the distortion gradient; both are multiplied
// by loss_scale (times 2^ceil(log2(info[0])) when scale_by_rays) and added to
// the density output's row 0, which is then scaled by t^2 where t < 1. Writes
// dfeat [32][feat_stride] fp32 and adds (accumulate=1) or stores the weight
// gradient into grad_net[9408] (fp32).
extern "C" int tn_mlp_backward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
                               uint32_t feat_stride, const void* params, uint32_t net_offset, const void* features,
                               const float* dir, const float* sample_t, const void* dL_drgb, uint32_t grad_stride,
                               const float* dL_ds, const float* dL_ds_extra, float loss_scale, int scale_by_rays,
                               void* dfeat, float* grad_net, int accumulate, void* workspace, size_t workspace_bytes,
                               const tn_appearance_t* appearance, hipStream_t stream) {
    TN_CHECK_ARG(params && features && dir && sample_t && dL_drgb && dL_ds && dfeat && grad_net && workspace);
    const tn_appearance_t app = appearance ? *appearance : tn_appearance_t{};
    TN_CHECK_ARG(!app.bias || (app.n_appearances > 0 && (reinterpret_cast<uintptr_t>(app.bias) & 15) == 0));
    TN_CHECK_ARG(!app.dz || app.bias);
    TN_CHECK_ARG(!scale_by_rays || info);
    TN_CHECK_ARG(n_max <= stride && n_max <= feat_stride && n_max <= grad_stride);
    TN_CHECK_ARG(128ull * feat_stride < (1ull << 32) && 3ull * stride < (1ull << 32) && 3ull * grad_stride < (1ull << 32));
    TN_CHECK_ARG(!app.dz || 256ull * stride < (1ull << 32));
    if (workspace_bytes < tn_mlp_backward_workspace_bytes(n_max)) return TN_ERR_OUT_OF_MEMORY;
    if (n_max == 0) return TN_OK;
    const unsigned blocks = mlp_blocks(n_max, (unsigned)tune_value(TN_TUNE_MLP_BWD_BLOCKS_PER_CU));
    float* partial = static_cast<float*>(workspace);
    if (precision == 1) {
        using P = PrecF16;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        if (!allow_lds(mlp_backward_kernel<P, float, P::T>, bwd_lds_bytes<P>())) return TN_ERR_LAUNCH;
        mlp_backward_kernel<P, float, P::T><<<blocks, 64 * kMlpWaves, bwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const float*>(features), dir, sample_t,
            static_cast<const P::T*>(dL_drgb), grad_stride, dL_ds, dL_ds_extra, loss_scale, scale_by_rays,
            static_cast<float*>(dfeat), partial, app);
    } else if (precision == 0) {
        using P = PrecF32;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        if (!allow_lds(mlp_backward_kernel<P, P::T, P::T>, bwd_lds_bytes<P>())) return TN_ERR_LAUNCH;
        mlp_backward_kernel<P, P::T, P::T><<<blocks, 64 * kMlpWaves, bwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, sample_t,
            static_cast<const P::T*>(dL_drgb), grad_stride, dL_ds, dL_ds_extra, loss_scale, scale_by_rays,
            static_cast<P::T*>(dfeat), partial, app);
    } else {
        return TN_ERR_INVALID_ARGUMENT;
    }
    reduce_partials_kernel<<<blocks_for(kNetParams, 64), 64 * kReduceWaves, 0, stream>>>(blocks, partial, grad_net,
                                                                                   accumulate);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_mlp_lds_bytes(int precision, int backward) {
    if (precision == 1) return backward ? bwd_lds_bytes<PrecF16>() : fwd_lds_bytes<PrecF16>();
    return backward ? bwd_lds_bytes<PrecF32>() : fwd_lds_bytes<PrecF32>();
}
