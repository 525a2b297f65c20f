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

// Activations of one 32-sample tile (the forward pass; the backward pass
// recomputes it instead of storing activations).
template <class P>
struct Tile {
    typename P::Op x[32 / P::KC];   // hash features, natural order
    typename P::Op sh[16 / P::KC];  // SH coefficients, natural order
    f16v h1[2], s, h2[2], h3[2], o;
};

template <class P, typename TF>
TN_DEVICE void load_inputs(Tile<P>& tl, bool valid, uint64_t sample, int h, const TF* __restrict__ feat,
                           uint32_t FS, const float* __restrict__ dir, uint32_t B, bool need_sh) {
#pragma unroll
    for (int ks = 0; ks < 32 / P::KC; ++ks) {
        float v[P::SL];
#pragma unroll
        for (int j = 0; j < P::SL; ++j) {
            const int row = nat_in<P>(ks, h, j);
            v[j] = valid ? (float)feat[(uint64_t)row * FS + sample] : 0.0f;
        }
        tl.x[ks] = P::from(v);
    }
    if (need_sh) {
        float shv[16];
        if (valid) sh4(dir[sample], dir[sample + B], dir[sample + 2 * (uint64_t)B], shv);
        else
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

template <class P>
TN_DEVICE void forward_tile(Tile<P>& tl, const FwdLds<P>& W, int lane, bool color) {
    using N = Net<P>;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        f16v acc = zero16();
#pragma unroll
        for (int ks = 0; ks < N::L1.n_ks; ++ks) acc = P::mma(load_a<P>(W.l1, N::L1, mt, ks, lane), tl.x[ks], acc);
        tl.h1[mt] = relu(acc);
    }
    tl.s = gemm_dt<P, 2>(W.l2, N::L2, 0, tl.h1, lane, N::L2.n_ks);
    if (!color) return;
    constexpr int half = kKsPerTile<P> / 2;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
        f16v acc = zero16();
#pragma unroll
        for (int ks = 0; ks < half; ++ks) acc = P::mma(load_a<P>(W.l3, N::L3, mt, ks, lane), dt_op<P>(tl.s, ks), acc);
#pragma unroll
        for (int ks = 0; ks < 16 / P::KC; ++ks)
            acc = P::mma(load_a<P>(W.l3, N::L3, mt, half + ks, lane), tl.sh[ks], acc);
        tl.h2[mt] = relu(acc);
    }
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) tl.h3[mt] = relu(gemm_dt<P, 2>(W.l4, N::L4, mt, tl.h2, lane, N::L4.n_ks));
    tl.o = gemm_dt<P, 2>(W.l5, N::L5, 0, tl.h3, lane, N::L5.n_ks);
}

TN_DEVICE float sigmoidf(float x) { return 1.0f / (1.0f + __expf(-x)); }

constexpr int kMlpWaves = 4;

template <class P, typename TF>
__global__ void __launch_bounds__(64 * kMlpWaves)
mlp_forward_kernel(uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t FS,
                   const typename P::T* __restrict__ params, const TF* __restrict__ feat,
                   const float* __restrict__ dir, int color, typename P::T* __restrict__ density_out,
                   typename P::T* __restrict__ rgb_out, uint32_t OS) {
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
        const uint64_t sample = (uint64_t)t * 32 + (lane & 31);
        const bool valid = sample < n;
        Tile<P> tl;
        load_inputs<P>(tl, valid, sample, h, feat, FS, dir, B, color != 0);
        forward_tile<P>(tl, W, lane, color != 0);
        if (valid && h == 0) {
            density_out[sample] = (typename P::T)tl.s[0];
            if (color) {
                rgb_out[sample] = (typename P::T)sigmoidf(tl.o[0]);
                rgb_out[sample + OS] = (typename P::T)sigmoidf(tl.o[1]);
                rgb_out[sample + 2 * (uint64_t)OS] = (typename P::T)sigmoidf(tl.o[2]);
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

// LDS transpose buffers: row stride 40 elements in mixed precision, so that
// a row's 8 consecutive samples (one fp16 MFMA operand) are one aligned
// 16-byte read; 33 in fp32 (scalar operands; the +1 spreads banks and the
// fp32 weights leave no room for more)
template <class P>
constexpr int kTW = P::SL == 8 ? 40 : 33;

// write an accumulator tile as rows [32*t, 32*t+32) x 32 samples
template <class P>
TN_DEVICE void put_tile(typename P::T* buf, int t, const f16v& D, int lane) {
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
        buf[row * kTW<P> + n] = (typename P::T)D[r];
    }
}

// write natural-order operands rows [base, base + 32/P::KC*P::KC)
template <class P, int NKS>
TN_DEVICE void put_nat(typename P::T* buf, int base, const typename P::Op (&x)[NKS], int lane) {
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks)
#pragma unroll
        for (int j = 0; j < P::SL; ++j) {
            const int row = base + nat_in<P>(ks, h, j);
            if constexpr (P::SL == 1) buf[row * kTW<P> + n] = x[ks];
            else buf[row * kTW<P> + n] = x[ks][j];
        }
}

// operand holding samples [ks*KC + h*SL, +SL) of row 32*t + (lane & 31)
template <class P>
TN_DEVICE typename P::Op get_rows(const typename P::T* buf, int t, int ks, int lane) {
    const int row = 32 * t + (lane & 31), h = lane >> 5;
    return *reinterpret_cast<const typename P::Op*>(buf + row * kTW<P> + ks * P::KC + h * P::SL);
}

TN_DEVICE void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// dW[mt][nt] += delta rows (tile mt) x act rows (tile nt), samples as K
template <class P, int MT, int NT>
TN_DEVICE void accumulate_dw(f16v (&dw)[MT][NT], const typename P::T* dbuf, const typename P::T* abuf, int lane) {
#pragma unroll
    for (int ks = 0; ks < 32 / P::KC; ++ks) {
        typename P::Op b[NT];
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) b[nt] = get_rows<P>(abuf, nt, ks, lane);
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            const typename P::Op a = get_rows<P>(dbuf, mt, ks, lane);
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) dw[mt][nt] = P::mma(a, b[nt], dw[mt][nt]);
        }
    }
}

TN_DEVICE f16v relu_mask(f16v d, const f16v& act) {
#pragma unroll
    for (int r = 0; r < 16; ++r) d[r] = act[r] > 0.0f ? d[r] : 0.0f;
    return d;
}

// An activation tile kept for the backward pass in the network precision
// (half the registers of the fp32 accumulator tile in mixed precision).
template <class P>
struct ActTile {
    typename P::T v[16];
};

template <class P>
TN_DEVICE ActTile<P> keep(const f16v& D) {
    ActTile<P> a;
#pragma unroll
    for (int r = 0; r < 16; ++r) a.v[r] = (typename P::T)D[r];
    return a;
}

template <class P>
TN_DEVICE f16v relu_mask(f16v d, const ActTile<P>& act) {
#pragma unroll
    for (int r = 0; r < 16; ++r) d[r] = act.v[r] > (typename P::T)0.0f ? d[r] : 0.0f;
    return d;
}

template <class P>
TN_DEVICE void put_tile(typename P::T* buf, int t, const ActTile<P>& A, int lane) {
    const int n = lane & 31, h = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 32 * t + 8 * (r >> 2) + 4 * h + (r & 3);
        buf[row * kTW<P> + n] = A.v[r];
    }
}

// add tile C (rows 32*mt.., cols 32*nt..) of a [n_out][n_in] gradient into LDS
TN_DEVICE void add_dw_tile(float* acc, const f16v& C, int mt, int nt, int n_out, int n_in, int lane) {
    const int col = 32 * nt + (lane & 31), h = lane >> 5;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 32 * mt + 8 * (r >> 2) + 4 * h + (r & 3);
        if (row < n_out && col < n_in) acc[row * n_in + col] += C[r];
    }
}

template <class P, typename TF, typename TG>
__global__ void __launch_bounds__(64 * kMlpWaves)
mlp_backward_kernel(uint32_t n_max, const uint32_t* __restrict__ info, uint32_t B, uint32_t FS,
                    const typename P::T* __restrict__ params, const TF* __restrict__ feat,
                    const float* __restrict__ dir, const float* __restrict__ sample_t,
                    const TG* __restrict__ dL_drgb, uint32_t GS, const float* __restrict__ dL_ds,
                    const float* __restrict__ dL_ds_extra, float loss_scale, TF* __restrict__ dfeat,
                    float* __restrict__ dw_partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    using T = typename P::T;
    using N = Net<P>;
    T* lds = reinterpret_cast<T*>(smem_raw);
    const FwdLds<P> W = stage_forward<P>(lds, params);
    const BwdLds<P> WB = stage_backward<P>(lds + N::kFwdElems, params);
    const int wid = threadIdx.x >> 6;
    T* tbuf = lds + N::kFwdElems + N::kBwdElems + wid * (2 * 64 * kTW<P>);
    T* dbuf = tbuf;               // deltas, up to 64 rows
    T* abuf = tbuf + 64 * kTW<P>;    // activations, up to 64 rows
    __syncthreads();

    const uint32_t n = info ? min(info[1], n_max) : n_max;
    const uint32_t n_tiles = (n + 31) / 32;
    const int lane = (int)__lane_id();
    const int h = lane >> 5;
    const uint32_t wave = blockIdx.x * kMlpWaves + wid;
    const uint32_t n_waves = gridDim.x * kMlpWaves;

    f16v dW1[2][1], dW2[1][2], dW3[2][1], dW4[2][2], dW5[1][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        dW1[a][0] = zero16();
        dW3[a][0] = zero16();
        dW2[0][a] = zero16();
        dW5[0][a] = zero16();
        dW4[a][0] = zero16();
        dW4[a][1] = zero16();
    }

    for (uint32_t t = wave; t < n_tiles; t += n_waves) {
        const uint64_t sample = (uint64_t)t * 32 + (lane & 31);
        const bool valid = sample < n;
        Tile<P> tl;
        load_inputs<P>(tl, valid, sample, h, feat, FS, dir, B, true);
        forward_tile<P>(tl, W, lane, true);
        // activations the backward pass needs, in the network precision
        const ActTile<P> a_h1[2] = {keep<P>(tl.h1[0]), keep<P>(tl.h1[1])};
        const ActTile<P> a_s = keep<P>(tl.s);
        const ActTile<P> a_h2[2] = {keep<P>(tl.h2[0]), keep<P>(tl.h2[1])};
        const ActTile<P> a_h3[2] = {keep<P>(tl.h3[0]), keep<P>(tl.h3[1])};

        // dL/do through the sigmoid (rows 0..2 live in lanes 0..31, registers 0..2)
        f16v dO = zero16();
        if (valid && h == 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float y = sigmoidf(tl.o[c]);
                dO[c] = (float)dL_drgb[sample + (uint64_t)c * GS] * y * (1.0f - y);
            }
        }
        // dH3 = relu'(h3) (W5^T dO); dW5 += dO h3^T
        f16v dH3[2];
        f16v dOt[1] = {dO};
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) dH3[mt] = relu_mask<P>(gemm_dt<P, 1>(WB.b5, N::B5, mt, dOt, lane, N::B5.n_ks), a_h3[mt]);
        put_tile<P>(dbuf, 0, dO, lane);
        put_tile<P>(abuf, 0, a_h3[0], lane);
        put_tile<P>(abuf, 1, a_h3[1], lane);
        wave_sync();
        accumulate_dw<P, 1, 2>(dW5, dbuf, abuf, lane);
        wave_sync();
        // dH2 = relu'(h2) (W4^T dH3); dW4 += dH3 h2^T
        f16v dH2[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) dH2[mt] = relu_mask<P>(gemm_dt<P, 2>(WB.b4, N::B4, mt, dH3, lane, N::B4.n_ks), a_h2[mt]);
        put_tile<P>(dbuf, 0, dH3[0], lane);
        put_tile<P>(dbuf, 1, dH3[1], lane);
        put_tile<P>(abuf, 0, a_h2[0], lane);
        put_tile<P>(abuf, 1, a_h2[1], lane);
        wave_sync();
        accumulate_dw<P, 2, 2>(dW4, dbuf, abuf, lane);
        wave_sync();
        // dXc = W3^T dH2 (rows 0..15 = d density outputs); dW3 += dH2 xc^T
        f16v dXc = gemm_dt<P, 2>(WB.b3, N::B3, 0, dH2, lane, N::B3.n_ks);
        put_tile<P>(dbuf, 0, dH2[0], lane);
        put_tile<P>(dbuf, 1, dH2[1], lane);
        put_tile<P>(abuf, 0, a_s, lane);  // rows 0..15 = density outputs (rows 16..31 overwritten next)
        put_nat<P, 16 / P::KC>(abuf, 16, tl.sh, lane);
        wave_sync();
        accumulate_dw<P, 2, 1>(dW3, dbuf, abuf, lane);
        wave_sync();
        // density-output gradient: row 0 also receives the loss-scaled
        // compositing (and distortion) gradient, then the whole row is scaled
        // by t^2 where t < 1 (reference nerf-network.cu:734-758)
        f16v dS = zero16();
#pragma unroll
        for (int r = 0; r < 8; ++r) dS[r] = dXc[r];  // rows < 16: registers 0..7
        if (valid && h == 0) {
            float g0 = dS[0] + loss_scale * dL_ds[sample];
            if (dL_ds_extra) g0 += loss_scale * dL_ds_extra[sample];
            const float tt = sample_t[sample];
            if (tt < 1.0f) g0 *= tt * tt;
            dS[0] = g0;
        }
        // dH1 = relu'(h1) (W2^T dS); dW2 += dS h1^T
        f16v dSt[1] = {dS};
        f16v dH1[2];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) dH1[mt] = relu_mask<P>(gemm_dt<P, 1>(WB.b2, N::B2, mt, dSt, lane, N::B2.n_ks), a_h1[mt]);
        put_tile<P>(dbuf, 0, dS, lane);
        put_tile<P>(abuf, 0, a_h1[0], lane);
        put_tile<P>(abuf, 1, a_h1[1], lane);
        wave_sync();
        accumulate_dw<P, 1, 2>(dW2, dbuf, abuf, lane);
        wave_sync();
        // dX = W1^T dH1 -> feature gradients; dW1 += dH1 x^T
        const f16v dX = gemm_dt<P, 2>(WB.b1, N::B1, 0, dH1, lane, N::B1.n_ks);
        if (valid) {
            const int nn = lane & 31;
            (void)nn;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * h + (r & 3);
                dfeat[(uint64_t)row * FS + sample] = (TF)dX[r];
            }
        }
        put_tile<P>(dbuf, 0, dH1[0], lane);
        put_tile<P>(dbuf, 1, dH1[1], lane);
        put_nat<P, 32 / P::KC>(abuf, 0, tl.x, lane);
        wave_sync();
        accumulate_dw<P, 2, 1>(dW1, dbuf, abuf, lane);
        wave_sync();
    }

    // workgroup reduction of the weight gradients (waves in turn, so the sum
    // order is fixed), then one partial per workgroup
    __syncthreads();
    float* red = reinterpret_cast<float*>(smem_raw);
    for (int e = threadIdx.x; e < kNetParams; e += blockDim.x) red[e] = 0.0f;
    __syncthreads();
    for (int w = 0; w < kMlpWaves; ++w) {
        if (wid == w) {
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                add_dw_tile(red + kW1, dW1[a][0], a, 0, 64, 32, lane);
                add_dw_tile(red + kW2, dW2[0][a], 0, a, 16, 64, lane);
                add_dw_tile(red + kW3, dW3[a][0], a, 0, 64, 32, lane);
                add_dw_tile(red + kW4, dW4[a][0], a, 0, 64, 64, lane);
                add_dw_tile(red + kW4, dW4[a][1], a, 1, 64, 64, lane);
                add_dw_tile(red + kW5, dW5[0][a], 0, a, 3, 64, lane);
            }
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
    if (e < (uint32_t)kNetParams)
        for (uint32_t p = w; p < n_part; p += kReduceWaves) s += partial[(uint64_t)p * kNetParams + e];
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
    const size_t t = (size_t)kMlpWaves * 2 * 64 * kTW<P> * sizeof(typename P::T);
    const size_t r = (size_t)kNetParams * sizeof(float);
    return (w + t) > r ? (w + t) : r;
}


static_assert(bwd_lds_bytes<PrecF16>() <= 160 * 1024 && bwd_lds_bytes<PrecF32>() <= 160 * 1024,
              "the backward kernel's LDS must fit one CU (160 KiB)");

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

// Persistent launches: as many workgroups as fit on the device at once (the
// backward kernel holds one workgroup per CU, the forward two), each looping
// over 32-sample tiles. Fewer workgroups also means fewer weight stagings and
// fewer gradient partials to reduce.
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

inline unsigned mlp_blocks(uint32_t n_max, unsigned per_cu) {
    const uint32_t tiles = (n_max + 31) / 32;
    const unsigned b = blocks_for(tiles, kMlpWaves);
    const unsigned cap = per_cu * compute_units();
    return b > cap ? cap : (b == 0 ? 1 : b);
}
constexpr unsigned kFwdBlocksPerCU = 2, kBwdBlocksPerCU = 1;

}  // namespace tn

using namespace tn;

// precision: 0 = fp32 (MFMA fp32), 1 = fp16 (MFMA f16, fp32 accumulate).
// params: the network's parameter buffer in that precision (hash table first,
// `net_offset` = its length). features [32][feat_stride] in that precision,
// dir [3][stride] fp32 (0.5*d + 0.5). Outputs: density (raw s) [stride] and
// rgb [3][out_stride] in that precision.
extern "C" int tn_mlp_forward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
                              uint32_t feat_stride, uint32_t out_stride, const void* params, uint32_t net_offset,
                              const void* features, const float* dir, int color, void* density_out, void* rgb_out,
                              hipStream_t stream) {
    TN_CHECK_ARG(params && features && density_out && (!color || (dir && rgb_out)));
    TN_CHECK_ARG(n_max <= stride && n_max <= feat_stride && n_max <= out_stride);
    if (n_max == 0) return TN_OK;
    const unsigned blocks = mlp_blocks(n_max, kFwdBlocksPerCU);
    if (precision == 1) {
        using P = PrecF16;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        mlp_forward_kernel<P, P::T><<<blocks, 64 * kMlpWaves, fwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, color,
            static_cast<P::T*>(density_out), static_cast<P::T*>(rgb_out), out_stride);
    } else if (precision == 0) {
        using P = PrecF32;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        mlp_forward_kernel<P, P::T><<<blocks, 64 * kMlpWaves, fwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, color,
            static_cast<P::T*>(density_out), static_cast<P::T*>(rgb_out), out_stride);
    } else {
        return TN_ERR_INVALID_ARGUMENT;
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" size_t tn_mlp_backward_workspace_bytes(uint32_t n_max) {
    return (size_t)mlp_blocks(n_max, kBwdBlocksPerCU) * kNetParams * sizeof(float);
}

// Backward of both MLPs for one batch. dL_drgb [3][grad_stride] (loss-scaled,
// in the network precision), dL_ds [stride] fp32 compositing gradient,
// dL_ds_extra (may be null) e.g. the distortion gradient; both are multiplied
// by loss_scale and added to the density output's row 0, which is then scaled
// by t^2 where t < 1. Writes dfeat [32][feat_stride] and adds (accumulate=1)
// or stores the weight gradient into grad_net[9408] (fp32).
extern "C" int tn_mlp_backward(int precision, uint32_t n_max, const uint32_t* info, uint32_t stride,
                               uint32_t feat_stride, const void* params, uint32_t net_offset, const void* features,
                               const float* dir, const float* sample_t, const void* dL_drgb, uint32_t grad_stride,
                               const float* dL_ds, const float* dL_ds_extra, float loss_scale, void* dfeat,
                               float* grad_net, int accumulate, void* workspace, size_t workspace_bytes,
                               hipStream_t stream) {
    TN_CHECK_ARG(params && features && dir && sample_t && dL_drgb && dL_ds && dfeat && grad_net && workspace);
    TN_CHECK_ARG(n_max <= stride && n_max <= feat_stride && n_max <= grad_stride);
    if (workspace_bytes < tn_mlp_backward_workspace_bytes(n_max)) return TN_ERR_OUT_OF_MEMORY;
    if (n_max == 0) return TN_OK;
    const unsigned blocks = mlp_blocks(n_max, kBwdBlocksPerCU);
    float* partial = static_cast<float*>(workspace);
    if (precision == 1) {
        using P = PrecF16;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        if (!allow_lds(mlp_backward_kernel<P, P::T, P::T>, bwd_lds_bytes<P>())) return TN_ERR_LAUNCH;
        mlp_backward_kernel<P, P::T, P::T><<<blocks, 64 * kMlpWaves, bwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, sample_t,
            static_cast<const P::T*>(dL_drgb), grad_stride, dL_ds, dL_ds_extra, loss_scale,
            static_cast<P::T*>(dfeat), partial);
    } else if (precision == 0) {
        using P = PrecF32;
        const auto* w = static_cast<const P::T*>(params) + net_offset;
        if (!allow_lds(mlp_backward_kernel<P, P::T, P::T>, bwd_lds_bytes<P>())) return TN_ERR_LAUNCH;
        mlp_backward_kernel<P, P::T, P::T><<<blocks, 64 * kMlpWaves, bwd_lds_bytes<P>(), stream>>>(
            n_max, info, stride, feat_stride, w, static_cast<const P::T*>(features), dir, sample_t,
            static_cast<const P::T*>(dL_drgb), grad_stride, dL_ds, dL_ds_extra, loss_scale,
            static_cast<P::T*>(dfeat), partial);
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
