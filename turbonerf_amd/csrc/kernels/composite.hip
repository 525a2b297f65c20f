// Volume-rendering compositing with random-background blending, smooth-L1
// loss and its backward pass (reference src/utils/training-kernels-fused.cuh:
// 227-498, src/core/nerf-network.cu:547-603).
//
// Training compositing runs in three launches, none of which walks a long ray
// serially. The batch is cut into chunks of 64 samples, one per wave:
//  (1) per chunk, segmented scans keyed by ray form each sample's prefix of
//      sigma*dt from its segment's start, and the segment sums of w c and w
//      with transmittance from that start. Rays that start and end in the
//      chunk are complete: their loss and dL/dRc are written here. For the
//      segments of lane 0 and lane 63 (rays entering or leaving the chunk)
//      the sums go to a per-chunk record;
//  (2) per ray crossing chunks, its records are chained: the sigma*dt prefix
//      at each chunk start, the totals R = sum w c and A = sum w (a segment's
//      sums scaled by the transmittance at its start), the loss and dL/dRc,
//      and the prefix of w g at each chunk start (from the same segment sums,
//      since g is linear in c). If transmittance falls below min_T inside a
//      segment, that segment is summed sample by sample up to the stop and
//      later segments count nothing, as in the reference's loop;
//  (3) per chunk, the prefixes are recomputed with the chained carry-ins and
//      give, for every sample i of a ray,
//        g_i       = sum_ch G_ch (c_i,ch A + R_ch - bg_ch)
//        dL/dsig_i = dt_i (T_{i+1} g_i - sum_{k>i} w_k g_k)
//        dL/dc_i   = scale G w_i A
//      with the suffix sum taken as the closed-form total A sum_ch
//      G_ch(2R_ch-bg_ch) minus the compensated inclusive prefix. This is the
//      reference's adjoint (its running "cumsum" subtraction) written for a
//      lane-parallel scan. Samples after the early-termination point
//      (T < min_T) get zero gradient, as in the reference.
// Forward-only compositing (rendering) walks each wave's window of rays.

#include <mutex>

#include "segscan.hpp"

namespace tn {

template <typename T>
TN_DEVICE float ldf(const T* p, uint64_t i) { return (float)p[i]; }

template <typename T>
TN_DEVICE void stf(T* p, uint64_t i, float v) { p[i] = (T)v; }

constexpr int kCompositeWaves = 4;  // waves per workgroup

TN_DEVICE float smooth_l1(float x) {
    const float ax = fabsf(x);
    return ax < 1.0f ? 0.5f * x * x : ax - 0.5f;
}
TN_DEVICE float smooth_l1_grad(float x) { return fabsf(x) < 1.0f ? x : copysignf(1.0f, x); }

// per-ray scratch rows (stride = ray capacity): A, R[3], G[3]
constexpr int kRayScratchRows = 7;

template <typename R>
struct RayGrad {
    R GR, GG, GB;  // dL/dRc per channel
};

// loss and dL/dRc of one ray from its totals (reference :430-477); the
// per-ray totals and dL/dRc go to scratch in the accumulation precision R
// (stored only when `store`; dL/dRc is returned either way)
template <typename R>
TN_DEVICE RayGrad<R> ray_loss(uint32_t r, uint32_t B, R rr, R rg, R rb, R ra, const float* __restrict__ bg,
                              const float* __restrict__ target, R inv_n_rays, float* __restrict__ loss,
                              R* __restrict__ scratch, uint32_t SB, bool store = true) {
    const R gr = target[r], gg = target[r + B], gb = target[r + 2 * B], ga = target[r + 3 * B];
    const R br = bg[r], bgg = bg[r + B], bb = bg[r + 2 * B];
    const R gac = (R)1 - ga;
    const R grc = gr * ga + br * gac, ggc = gg * ga + bgg * gac, gbc = gb * ga + bb * gac;
    const R rac = (R)1 - ra;
    const R rrc = rr * ra + br * rac, rgc = rg * ra + bgg * rac, rbc = rb * ra + bb * rac;
    const R dr = rrc - grc, dg = rgc - ggc, db = rbc - gbc, da = rac - gac;
    auto sl1 = [](R x) { const R ax = x < (R)0 ? -x : x; return ax < (R)1 ? (R)0.5 * x * x : ax - (R)0.5; };
    auto sl1g = [](R x) { const R ax = x < (R)0 ? -x : x; return ax < (R)1 ? x : (x < (R)0 ? (R)-1 : (R)1); };
    const RayGrad<R> G = {inv_n_rays * sl1g(dr), inv_n_rays * sl1g(dg), inv_n_rays * sl1g(db)};
    if (!store) return G;
    loss[r] = (float)sl1(dr);
    loss[r + B] = (float)sl1(dg);
    loss[r + 2 * B] = (float)sl1(db);
    loss[r + 3 * B] = (float)sl1(da);
    if (scratch) {
        scratch[r] = ra;
        scratch[r + SB] = rr;
        scratch[r + 2 * SB] = rg;
        scratch[r + 3 * SB] = rb;
        scratch[r + 4 * SB] = G.GR;
        scratch[r + 5 * SB] = G.GG;
        scratch[r + 6 * SB] = G.GB;
    }
    return G;
}

// Per-lane recomputation shared by both passes: transmittance from the
// segmented prefix of sigma*dt, alpha, weight and the early-stop mask.
template <typename R>
struct SampleState {
    R sigma, dt, T, T_next, alpha, w;
    Acc<R> P;  // inclusive prefix of sigma * dt along the ray
    bool inc;
};

// Per-sample inputs of one 64-sample chunk, loaded one chunk ahead of their
// use. Lanes past the window read the window's last sample (in bounds) and
// are masked where the values are used.
struct ChunkIn {
    float d, dt, cr, cg, cb;
};

template <typename TIn>
TN_DEVICE ChunkIn load_chunk(uint32_t c, uint32_t s1, uint32_t lane, const TIn* __restrict__ density,
                             const float* __restrict__ dtb, const TIn* __restrict__ color, uint32_t CB) {
    const uint32_t s = c + lane;
    const uint32_t sl = s < s1 ? s : s1 - 1;
    ChunkIn x;
    x.d = ldf(density, sl);
    x.dt = dtb[sl];
    x.cr = ldf(color, sl);
    x.cg = ldf(color, sl + CB);
    x.cb = ldf(color, sl + 2 * (uint64_t)CB);
    return x;
}

template <typename R>
TN_DEVICE SampleState<R> sample_state(bool valid, int ray, const ChunkIn& in, int& carry_ray, Acc<R>& carry_P,
                                      float min_T, bool raw_density = true) {
    SampleState<R> st;
    st.sigma = 0;
    st.dt = 0;
    if (valid) {
        st.sigma = raw_density ? r_exp<R>((R)fminf(in.d, 12.0f) - (R)1) : (R)in.d;
        st.dt = (R)in.dt;
    }
    const R tau = st.sigma * st.dt;
    Acc<R> P = Acc<R>::of(tau).scan(ray);
    if (ray == carry_ray) P = P + carry_P;
    const R p_incl = P.val();
    const R p_excl = p_incl - tau;
    st.T = r_exp<R>(-p_excl);
    st.T_next = r_exp<R>(-p_incl);
    st.alpha = r_alpha<R>(tau);
    st.w = st.T * st.alpha;
    st.inc = valid && st.T >= (R)min_T;
    st.P = P;
    // carry of the prefix for the ray continuing past lane 63
    const int ray63 = __builtin_amdgcn_readlane(ray, 63);
    carry_P = P.bcast(63);
    carry_ray = ray63;
    return st;
}

// per chunk: two segment records of kChunkRec sums (sigma*dt, w c_r, w c_g,
// w c_b, w) and a carry-in pair (sigma*dt prefix, w g prefix)
constexpr int kChunkRec = 5;

template <typename R>
TN_DEVICE void put_rec(Acc<R>* __restrict__ dst, const Acc<R>& P, const Acc<R>& sr, const Acc<R>& sg,
                       const Acc<R>& sb, const Acc<R>& sa) {
    dst[0] = P;
    dst[1] = sr;
    dst[2] = sg;
    dst[3] = sb;
    dst[4] = sa;
}

// (1) chunk sums, complete rays' losses
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kCompositeWaves)
composite_chunk_kernel(const uint32_t* __restrict__ info, uint32_t B, uint32_t PB, uint32_t SB, float min_T,
                       const uint32_t* __restrict__ steps, const uint32_t* __restrict__ offs,
                       const float* __restrict__ bg, const float* __restrict__ target, const TIn* __restrict__ density,
                       const TIn* __restrict__ color, const float* __restrict__ dtb, float* __restrict__ loss,
                       R* __restrict__ scratch, Acc<R>* __restrict__ rec) {
    __shared__ int slots[kCompositeWaves][128];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    const uint32_t q = blockIdx.x * kCompositeWaves + wid;
    const uint32_t n_rays = info[0], n_samples = info[1];
    const uint32_t n_chunks = (n_samples + 63) / 64;
    if (n_rays == 0 || q >= max(n_chunks, 1u)) return;
    const R inv_n_rays = (R)1 / (R)n_rays;
    int* slot = slots[wid];
    auto empty_ray = [&](uint32_t r) {
        ray_loss<R>(r, B, (R)0, (R)0, (R)0, (R)0, bg, target, inv_n_rays, loss, (R*)nullptr, SB);
    };
    if (n_chunks == 0) {  // every ray is empty
        for (uint32_t r = lane; r < n_rays; r += 64) empty_ray(r);
        return;
    }
    const uint32_t c = 64 * q;
    RayWalk wk = chunk_walk(c, n_rays, offs, steps);
    const int carried = wk.carry_ray;
    const uint64_t s = (uint64_t)c + lane;
    const bool valid = s < n_samples;
    const ChunkIn in = load_chunk(c, n_samples, lane, density, dtb, color, PB);
    const LaneRay lr = chunk_rays(c, n_rays, wk, offs, steps, slot, empty_ray);
    const int ray = valid ? lr.ray : -2 - (int)lane;  // unique keys keep padding lanes out of every segment
    int no_carry = -1;
    Acc<R> no_P = Acc<R>::zero();
    const SampleState<R> st = sample_state<R>(valid, ray, in, no_carry, no_P, min_T);
    // Rays starting in this chunk see their true transmittance and stop as
    // the reference does; the carried ray's segment is summed whole (its
    // stop, if any, is found per ray).
    const bool own = valid && ray != carried;
    const R wv = (own ? st.inc : valid) ? st.w : (R)0;
    const float cR = valid ? in.cr : 0.0f, cG = valid ? in.cg : 0.0f, cB = valid ? in.cb : 0.0f;
    const Acc<R> sr = Acc<R>::of(wv * (R)cR).scan(ray), sg = Acc<R>::of(wv * (R)cG).scan(ray);
    const Acc<R> sb = Acc<R>::of(wv * (R)cB).scan(ray), sa = Acc<R>::of(wv).scan(ray);
    const bool tail = valid && s + 1 == (uint64_t)lr.end;
    if (tail && own)
        ray_loss<R>((uint32_t)ray, B, sr.val(), sg.val(), sb.val(), sa.val(), bg, target, inv_n_rays, loss, scratch,
                    SB);
    // the records: lane 0's segment at its last lane, lane 63's segment
    const int ray_next = __shfl_down(ray, 1, 64);
    const bool seg_last = lane == 63 || ray_next != ray;
    Acc<R>* rq = rec + (uint64_t)q * 2 * kChunkRec;
    if (valid && seg_last && ray == __builtin_amdgcn_readlane(ray, 0)) put_rec<R>(rq, st.P, sr, sg, sb, sa);
    if (valid && lane == 63) put_rec<R>(rq + kChunkRec, st.P, sr, sg, sb, sa);
    // zero-length rays after the last sample
    if (q + 1 == n_chunks)
        for (uint32_t r = wk.r_scan + lane; r < n_rays; r += 64) empty_ray(r);
}

// (2) one wave per ray spanning several chunks, a lane per chunk of the ray
// (64 at a time): the sigma*dt prefixes at the chunk starts by a lane scan,
// the stop (the first chunk whose end sees T < min_T: its samples are summed
// in order up to the stop by that lane), the totals, loss and dL/dRc, and the
// prefixes of w g at the chunk starts by a second lane scan
template <typename TIn, typename R>
TN_DEVICE void composite_one_ray(uint32_t r, const uint32_t* __restrict__ info, uint32_t B, uint32_t PB, uint32_t SB, float min_T,
                     const uint32_t* __restrict__ steps, const uint32_t* __restrict__ offs,
                     const float* __restrict__ bg, const float* __restrict__ target, const TIn* __restrict__ density,
                     const TIn* __restrict__ color, const float* __restrict__ dtb, float* __restrict__ loss,
                     R* __restrict__ scratch, const Acc<R>* __restrict__ rec, Acc<R>* __restrict__ carry) {
    const uint32_t lane = __lane_id();
    const uint32_t n_rays = info[0];
    if (r >= n_rays) return;
    const uint32_t o = offs[r], n = steps[r];
    if (n == 0) return;
    const uint32_t qs = o / 64, qe = (o + n - 1) / 64;
    if (qs == qe) return;  // complete in one chunk
    auto rec_of = [&](uint32_t k) { return rec + ((uint64_t)k * 2 + (k == qs ? 1u : 0u)) * kChunkRec; };
    // totals: lanes before the stop add their segment sums scaled by the
    // transmittance at the segment start, the stop lane its partial sums
    Acc<R> P0 = Acc<R>::zero();  // sigma*dt before this pass's first chunk
    Acc<R> tr = Acc<R>::zero(), tg = Acc<R>::zero(), tb = Acc<R>::zero(), ta = Acc<R>::zero();
    uint32_t q_stop = 0xFFFFFFFFu;  // chunk whose segment holds the stop, if any
    Acc<R> pr = Acc<R>::zero(), pg = Acc<R>::zero(), pb = Acc<R>::zero(), pa = Acc<R>::zero();  // stop lane's
    for (uint32_t k0 = qs; k0 <= qe; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool act = k <= qe;
        const Acc<R>* rk = rec_of(act ? k : qe);
        const Acc<R> Pk = act ? rk[0] : Acc<R>::zero();
        const Acc<R> Pin = P0 + acc_exclusive<R>(Pk);
        const Acc<R> Pout = Pin + Pk;
        if (act && k > qs) carry[(uint64_t)k * 2] = Pin;
        const uint64_t low = __ballot(act && !(r_exp<R>(-Pout.val()) >= (R)min_T));
        const uint32_t stop_lane = low ? (uint32_t)__builtin_ctzll(low) : 64u;
        Acc<R> cr = Acc<R>::zero(), cg = Acc<R>::zero(), cb = Acc<R>::zero(), ca = Acc<R>::zero();
        if (q_stop == 0xFFFFFFFFu) {
            if (act && lane < stop_lane) {
                const R Tin = r_exp<R>(-Pin.val());
                cr = acc_scaled<R>(rk[1], Tin);
                cg = acc_scaled<R>(rk[2], Tin);
                cb = acc_scaled<R>(rk[3], Tin);
                ca = acc_scaled<R>(rk[4], Tin);
            } else if (lane == stop_lane) {  // the stop lies in this segment: its samples in order
                const uint32_t s0 = max(o, 64 * k), s1 = min(o + n, 64 * k + 64);
                Acc<R> p = Pin;
                for (uint32_t sm = s0; sm < s1; ++sm) {
                    const R T = r_exp<R>(-p.val());
                    if (!(T >= (R)min_T)) break;
                    const R sigma = r_exp<R>((R)fminf(ldf(density, sm), 12.0f) - (R)1);
                    const R tau = sigma * (R)dtb[sm];
                    const R w = T * r_alpha<R>(tau);
                    pr = pr + Acc<R>::of(w * (R)ldf(color, sm));
                    pg = pg + Acc<R>::of(w * (R)ldf(color, sm + PB));
                    pb = pb + Acc<R>::of(w * (R)ldf(color, sm + 2 * (uint64_t)PB));
                    pa = pa + Acc<R>::of(w);
                    p = p + Acc<R>::of(tau);
                }
                cr = pr;
                cg = pg;
                cb = pb;
                ca = pa;
            }
            tr = tr + cr.scan(0).bcast(63);
            tg = tg + cg.scan(0).bcast(63);
            tb = tb + cb.scan(0).bcast(63);
            ta = ta + ca.scan(0).bcast(63);
            if (stop_lane < 64) q_stop = k0 + stop_lane;
        }
        P0 = Pout.bcast(63);
    }
    // the stop lane's partial sums, for the w g prefixes
    if (q_stop != 0xFFFFFFFFu) {
        const int src = (int)((q_stop - qs) & 63u);
        pr = pr.bcast(src);
        pg = pg.bcast(src);
        pb = pb.bcast(src);
        pa = pa.bcast(src);
    }
    const R inv_n_rays = (R)1 / (R)n_rays;
    const R A = ta.val(), RR = tr.val(), RG = tg.val(), RB = tb.val();
    const RayGrad<R> G = ray_loss<R>(r, B, RR, RG, RB, A, bg, target, inv_n_rays, loss, scratch, SB, lane == 0);
    const R br = bg[r], bgg = bg[r + B], bb = bg[r + 2 * B];
    // prefix of w g at each chunk start: sum_seg w g = sum_ch G_ch (A sum w c_ch + (R_ch - bg_ch) sum w)
    auto wg_of = [&](R swr, R swg, R swb, R swa) {
        return Acc<R>::of(G.GR * A * swr) + Acc<R>::of(G.GR * (RR - br) * swa) + Acc<R>::of(G.GG * A * swg) +
               Acc<R>::of(G.GG * (RG - bgg) * swa) + Acc<R>::of(G.GB * A * swb) + Acc<R>::of(G.GB * (RB - bb) * swa);
    };
    Acc<R> Q0 = Acc<R>::zero();
    P0 = Acc<R>::zero();
    for (uint32_t k0 = qs; k0 <= qe; k0 += 64) {
        const uint32_t k = k0 + lane;
        const bool act = k <= qe;
        const Acc<R>* rk = rec_of(act ? k : qe);
        const Acc<R> Pk = act ? rk[0] : Acc<R>::zero();
        const Acc<R> Pin = P0 + acc_exclusive<R>(Pk);
        Acc<R> wg = Acc<R>::zero();
        if (act && k < q_stop) {
            const R Tin = r_exp<R>(-Pin.val());
            wg = wg_of(Tin * rk[1].val(), Tin * rk[2].val(), Tin * rk[3].val(), Tin * rk[4].val());
        } else if (act && k == q_stop) {
            wg = wg_of(pr.val(), pg.val(), pb.val(), pa.val());
        }
        const Acc<R> Qin = Q0 + acc_exclusive<R>(wg);
        if (act && k > qs) carry[(uint64_t)k * 2 + 1] = Qin;
        Q0 = (Qin + wg).bcast(63);
        P0 = (Pin + Pk).bcast(63);
    }
}

// rays r = wave, wave + the grid's waves, ...: the grid is capped (most of a
// training batch's ray capacity is unused), so idle workgroups are not launched
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kCompositeWaves)
composite_ray_kernel(const uint32_t* __restrict__ info, uint32_t B, uint32_t PB, uint32_t SB, float min_T,
                     const uint32_t* __restrict__ steps, const uint32_t* __restrict__ offs,
                     const float* __restrict__ bg, const float* __restrict__ target, const TIn* __restrict__ density,
                     const TIn* __restrict__ color, const float* __restrict__ dtb, float* __restrict__ loss,
                     R* __restrict__ scratch, const Acc<R>* __restrict__ rec, Acc<R>* __restrict__ carry) {
    const uint32_t n_rays = info[0];
    for (uint32_t r = blockIdx.x * kCompositeWaves + (threadIdx.x >> 6); r < n_rays; r += gridDim.x * kCompositeWaves)
        composite_one_ray<TIn, R>(r, info, B, PB, SB, min_T, steps, offs, bg, target, density, color, dtb, loss, scratch, rec, carry);
}

// (3) gradients per chunk with the chained carry-ins
template <typename TIn, typename TGrad, typename R>
__global__ void __launch_bounds__(64 * kCompositeWaves)
composite_grad_kernel(const uint32_t* __restrict__ info, uint32_t B, uint32_t PB, uint32_t SB, float grad_scale,
                      int scale_by_rays, float min_T, const uint32_t* __restrict__ steps, const uint32_t* __restrict__ offs,
                      const float* __restrict__ bg, const TIn* __restrict__ density, const TIn* __restrict__ color,
                      const float* __restrict__ dtb, float* __restrict__ dL_ddensity, TGrad* __restrict__ dL_dcolor,
                      const R* __restrict__ scratch, const Acc<R>* __restrict__ carry) {
    __shared__ int slots[kCompositeWaves][128];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    const uint32_t q = blockIdx.x * kCompositeWaves + wid;
    const uint32_t n_rays = info[0], n_samples = info[1];
    const uint32_t n_chunks = (n_samples + 63) / 64;
    if (n_rays == 0 || q >= n_chunks) return;
    int* slot = slots[wid];
    const uint32_t c = 64 * q;
    RayWalk wk = chunk_walk(c, n_rays, offs, steps);
    const int carried = wk.carry_ray;
    Acc<R> Pin = Acc<R>::zero(), Qin = Acc<R>::zero();
    if (carried >= 0) {
        Pin = carry[(uint64_t)q * 2];
        Qin = carry[(uint64_t)q * 2 + 1];
    }
    const uint64_t s = (uint64_t)c + lane;
    const bool valid = s < n_samples;
    const ChunkIn in = load_chunk(c, n_samples, lane, density, dtb, color, PB);
    auto no_op = [](uint32_t) {};
    const LaneRay lr = chunk_rays(c, n_rays, wk, offs, steps, slot, no_op);
    const int ray = valid ? lr.ray : -2 - (int)lane;
    int carry_ray = carried;
    Acc<R> carry_P = Pin;
    const SampleState<R> st = sample_state<R>(valid, ray, in, carry_ray, carry_P, min_T);
    R g = 0, wg = 0, A = 0, GR = 0, GG = 0, GB = 0, Wg_tot = 0;
    if (valid) {
        const uint32_t rr = (uint32_t)ray;
        A = scratch[rr];
        const R RR = scratch[rr + SB], RG = scratch[rr + 2 * SB], RB = scratch[rr + 3 * SB];
        GR = scratch[rr + 4 * SB];
        GG = scratch[rr + 5 * SB];
        GB = scratch[rr + 6 * SB];
        const R br = bg[rr], bgg = bg[rr + B], bb = bg[rr + 2 * B];
        const R cR = in.cr, cG = in.cg, cB = in.cb;
        g = GR * (cR * A + RR - br) + GG * (cG * A + RG - bgg) + GB * (cB * A + RB - bb);
        Wg_tot = A * (GR * ((R)2 * RR - br) + GG * ((R)2 * RG - bgg) + GB * ((R)2 * RB - bb));
        wg = st.inc ? st.w * g : (R)0;
    }
    Acc<R> Q = Acc<R>::of(wg).scan(ray);
    if (ray == carried) Q = Q + Qin;
    if (!valid) return;
    float dsig = 0.0f, gc_r = 0.0f, gc_g = 0.0f, gc_b = 0.0f;
    if (st.inc) {
        const R suffix = Wg_tot - Q.val();
        dsig = (float)(st.dt * (st.T_next * g - suffix) * st.sigma);
        const float gs = scale_by_rays ? grad_scale * ray_pow2(n_rays) : grad_scale;
        const R k = (R)gs * st.w * A;
        gc_r = (float)(k * GR);
        gc_g = (float)(k * GG);
        gc_b = (float)(k * GB);
    }
    dL_ddensity[s] = dsig;
    stf(dL_dcolor, s, gc_r);
    stf(dL_dcolor, s + PB, gc_g);
    stf(dL_dcolor, s + 2 * (uint64_t)PB, gc_b);
}

// Forward-only compositing: premultiplied rgb and alpha per ray, with the
// reference's early termination. `raw_density` selects sigma = exp(min(s,12)-1)
// of a network output, otherwise the input already is sigma. As in the
// reference's render compositing (rendering-kernels.cu:570-578, colour divided
// by alpha only when alpha > 1e-4, else zero), a sample's colour counts only
// when its alpha exceeds 1e-4; its alpha always counts.
template <typename TIn, typename R>
__global__ void __launch_bounds__(64 * kCompositeWaves)
composite_forward_kernel(const uint32_t* __restrict__ info, uint32_t B, uint32_t OB, uint32_t S, float min_T,
                         int raw_density, const uint32_t* __restrict__ steps, const uint32_t* __restrict__ offs,
                         const TIn* __restrict__ density, const TIn* __restrict__ color,
                         const float* __restrict__ dtb, float* __restrict__ rgba) {
    __shared__ int slots[kCompositeWaves][128];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t lane = __lane_id();
    const uint32_t w = blockIdx.x * kCompositeWaves + wid;
    const uint32_t n_rays = info[0], n_samples = info[1];
    const WaveRange wr = wave_range(w, S, n_rays, n_samples, offs);
    if (!wr.active) return;
    int* slot = slots[wid];
    auto empty_ray = [&](uint32_t r) {
        rgba[r] = 0.0f;
        rgba[r + OB] = 0.0f;
        rgba[r + 2 * OB] = 0.0f;
        rgba[r + 3 * OB] = 0.0f;
    };
    RayWalk wk;
    wk.r_scan = wr.r0;
    Acc<R> carry_P = Acc<R>::zero();
    Acc<R> cr = Acc<R>::zero(), cg = Acc<R>::zero(), cb = Acc<R>::zero(), ca = Acc<R>::zero();
    ChunkIn in = {};
    if (wr.s0 < wr.s1) in = load_chunk(wr.s0, wr.s1, lane, density, dtb, color, B);
    for (uint32_t c = wr.s0; c < wr.s1; c += 64) {
        const uint64_t s = (uint64_t)c + lane;
        const bool valid = s < wr.s1;
        const ChunkIn cur = in;
        if (c + 64 < wr.s1) in = load_chunk(c + 64, wr.s1, lane, density, dtb, color, B);
        const int prev_carry = wk.carry_ray;
        const LaneRay lr = chunk_rays(c, wr.r1, wk, offs, steps, slot, empty_ray);
        const int ray = valid ? lr.ray : -2 - (int)lane;
        SampleState<R> st = sample_state<R>(valid, ray, cur, wk.carry_ray, carry_P, min_T, raw_density != 0);
        const float cR = valid ? cur.cr : 0.0f, cG = valid ? cur.cg : 0.0f, cB = valid ? cur.cb : 0.0f;
        const R wv = st.inc ? st.w : (R)0;
        const R wc = st.alpha > (R)1e-4f ? wv : (R)0;
        Acc<R> sr = Acc<R>::of(wc * (R)cR).scan(ray), sg = Acc<R>::of(wc * (R)cG).scan(ray);
        Acc<R> sb = Acc<R>::of(wc * (R)cB).scan(ray), sa = Acc<R>::of(wv).scan(ray);
        if (ray == prev_carry) {
            sr = sr + cr;
            sg = sg + cg;
            sb = sb + cb;
            sa = sa + ca;
        }
        const bool tail = valid && s + 1 == (uint64_t)lr.end;
        if (tail) {
            const uint32_t rr = (uint32_t)ray;
            rgba[rr] = (float)sr.val();
            rgba[rr + OB] = (float)sg.val();
            rgba[rr + 2 * OB] = (float)sb.val();
            rgba[rr + 3 * OB] = (float)sa.val();
        }
        const bool tail63 = lane63((int)tail) != 0;
        cr = sr.bcast(63);
        cg = sg.bcast(63);
        cb = sb.bcast(63);
        ca = sa.bcast(63);
        wk.carry_end = lane63(lr.end);
        if (tail63 || !lane63((int)valid)) wk.carry_ray = -1;
    }
    for (uint32_t r = wk.r_scan + lane; r < wr.r1; r += 64) empty_ray(r);
}

// deterministic mean of the per-ray losses: (1/n_rays) sum_r sum_ch loss
__global__ void reduce_loss_kernel(const uint32_t* __restrict__ info, uint32_t B, uint32_t n_ch,
                                   const float* __restrict__ loss, float* __restrict__ out) {
    __shared__ double part[16];
    const uint32_t n = info[0];
    double acc = 0.0;
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) {
        float s = 0.0f;
        for (uint32_t k = 0; k < n_ch; ++k) s += loss[r + (uint64_t)k * B];
        acc += (double)s;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_down(acc, d, 64);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) t += part[k];
        out[0] = n > 0 ? (float)(t / (double)n) : 0.0f;
    }
}

// composite scratch: per-ray rows, then per-chunk records and carry-ins
inline size_t comp_align(size_t x) { return (x + 255) & ~(size_t)255; }
inline uint32_t comp_chunks(uint32_t sample_stride) { return sample_stride / 64 + 1; }
inline size_t comp_rays_bytes(uint32_t n_rays_max) { return comp_align((size_t)kRayScratchRows * n_rays_max * 8); }
inline size_t comp_rec_bytes(uint32_t sample_stride) {
    return comp_align((size_t)comp_chunks(sample_stride) * 2 * kChunkRec * 8);
}
inline size_t comp_carry_bytes(uint32_t sample_stride) {
    return comp_align((size_t)comp_chunks(sample_stride) * 2 * 8);
}
static_assert(sizeof(Acc<float>) == 8 && sizeof(Acc<double>) == 8, "8-byte accumulators");

// A training step's metrics in one launch: out[0] = mean over rays of the
// four loss channels, out[1] = mean distortion loss (0 without one),
// total[0] += info[1] when a counter is given, and ++step[0] when a step
// counter is given (the last kernel of a step).
__global__ void step_metrics_kernel(const uint32_t* __restrict__ info, uint32_t B, const float* __restrict__ loss,
                                    const float* __restrict__ dist_loss, float* __restrict__ out,
                                    unsigned long long* __restrict__ total, uint32_t* __restrict__ step) {
    __shared__ double part[2][16];
    const uint32_t n = info[0];
    double acc = 0.0, accd = 0.0;
    for (uint32_t r = threadIdx.x; r < n; r += blockDim.x) {
        float s = 0.0f;
        for (uint32_t k = 0; k < 4; ++k) s += loss[r + (uint64_t)k * B];
        acc += (double)s;
        if (dist_loss) accd += (double)dist_loss[r];
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        acc += __shfl_down(acc, d, 64);
        accd += __shfl_down(accd, d, 64);
    }
    if (__lane_id() == 0) {
        part[0][threadIdx.x >> 6] = acc;
        part[1][threadIdx.x >> 6] = accd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0, td = 0.0;
        for (uint32_t k = 0; k < blockDim.x / 64; ++k) {
            t += part[0][k];
            td += part[1][k];
        }
        out[0] = n > 0 ? (float)(t / (double)n) : 0.0f;
        out[1] = n > 0 && dist_loss ? (float)(td / (double)n) : 0.0f;
        if (total) *total += info[1];
        if (step) *step += 1u;
    }
}

template <typename TIn, typename TGrad, typename R>
int launch_composite(const uint32_t* info, uint32_t n_rays_max, uint32_t B, uint32_t PB, float grad_scale,
                     int scale_by_rays, float min_T, const uint32_t* steps, const uint32_t* offs, const float* bg, const float* target,
                     const void* density, const void* color, const float* dt, float* loss, float* dL_ddensity,
                     void* dL_dcolor, void* scratch, hipStream_t stream) {
    unsigned char* base = static_cast<unsigned char*>(scratch);
    R* rays = reinterpret_cast<R*>(base);
    Acc<R>* rec = reinterpret_cast<Acc<R>*>(base + comp_rays_bytes(n_rays_max));
    Acc<R>* carry = reinterpret_cast<Acc<R>*>(base + comp_rays_bytes(n_rays_max) + comp_rec_bytes(PB));
    const uint32_t chunk_blocks = (comp_chunks(PB) + kCompositeWaves - 1) / kCompositeWaves;
    const TIn* d = static_cast<const TIn*>(density);
    const TIn* c = static_cast<const TIn*>(color);
    composite_chunk_kernel<TIn, R><<<chunk_blocks, 64 * kCompositeWaves, 0, stream>>>(
        info, B, PB, n_rays_max, min_T, steps, offs, bg, target, d, c, dt, loss, rays, rec);
    composite_ray_kernel<TIn, R><<<ray_wave_blocks(n_rays_max, kCompositeWaves), 64 * kCompositeWaves, 0, stream>>>(
        info, B, PB, n_rays_max, min_T, steps, offs, bg, target, d, c, dt, loss, rays, rec, carry);
    composite_grad_kernel<TIn, TGrad, R><<<chunk_blocks, 64 * kCompositeWaves, 0, stream>>>(
        info, B, PB, n_rays_max, grad_scale, scale_by_rays, min_T, steps, offs, bg, d, c, dt, dL_ddensity,
        static_cast<TGrad*>(dL_dcolor), rays, carry);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

}  // namespace tn

using namespace tn;

extern "C" size_t tn_composite_scratch_bytes(uint32_t n_rays_max, uint32_t sample_stride) {
    return comp_rays_bytes(n_rays_max) + comp_rec_bytes(sample_stride) + comp_carry_bytes(sample_stride);
}

// dtype: 0 = fp32 network outputs and fp32 colour gradients,
//        1 = fp16 network outputs and fp16 colour gradients,
//        2 = fp32 in/out with fp64 accumulation (oracle mode)
extern "C" int tn_composite_train(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t stride,
                                  uint32_t sample_stride, float grad_scale, int scale_by_rays, float min_transmittance,
                                  const uint32_t* ray_steps,
                                  const uint32_t* ray_offset, const float* bg_rgb, const float* target_rgba,
                                  const void* density, const void* color, const float* dt, float* loss,
                                  float* dL_ddensity, void* dL_dcolor, void* ray_scratch, size_t scratch_bytes,
                                  hipStream_t stream) {
    TN_CHECK_ARG(info && ray_steps && ray_offset && bg_rgb && target_rgba && density && color && dt && loss &&
                 dL_ddensity && dL_dcolor && ray_scratch);
    TN_CHECK_ARG(n_rays_max > 0 && n_rays_max <= stride && sample_stride > 0 && sample_stride < (1u << 31));
    if (scratch_bytes < tn_composite_scratch_bytes(n_rays_max, sample_stride)) return TN_ERR_OUT_OF_MEMORY;
    TN_CHECK_ARG((reinterpret_cast<uintptr_t>(ray_scratch) & 255) == 0);
    switch (dtype) {
        case 0:
            return launch_composite<float, float, float>(info, n_rays_max, stride, sample_stride, grad_scale, scale_by_rays,
                                                         min_transmittance,
                                                         ray_steps, ray_offset, bg_rgb, target_rgba, density, color,
                                                         dt, loss, dL_ddensity, dL_dcolor, ray_scratch, stream);
        case 1:
            return launch_composite<_Float16, _Float16, float>(info, n_rays_max, stride, sample_stride, grad_scale,
                                                               scale_by_rays, min_transmittance, ray_steps, ray_offset, bg_rgb,
                                                               target_rgba, density, color, dt, loss, dL_ddensity,
                                                               dL_dcolor, ray_scratch, stream);
        case 2:
            return launch_composite<float, float, double>(info, n_rays_max, stride, sample_stride, grad_scale,
                                                          scale_by_rays, min_transmittance, ray_steps, ray_offset, bg_rgb,
                                                          target_rgba, density, color, dt, loss, dL_ddensity,
                                                          dL_dcolor, ray_scratch, stream);
        default:
            return TN_ERR_INVALID_ARGUMENT;
    }
}

extern "C" int tn_reduce_loss(const uint32_t* info, uint32_t stride, uint32_t n_channels, const float* loss,
                              float* out, hipStream_t stream) {
    TN_CHECK_ARG(info && loss && out && n_channels >= 1 && n_channels <= 4);
    reduce_loss_kernel<<<1, 1024, 0, stream>>>(info, stride, n_channels, loss, out);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

extern "C" int tn_step_metrics(const uint32_t* info, uint32_t stride, const float* loss, const float* dist_loss,
                               float* out, unsigned long long* samples_total, uint32_t* step_counter,
                               hipStream_t stream) {
    TN_CHECK_ARG(info && loss && out);
    step_metrics_kernel<<<1, 1024, 0, stream>>>(info, stride, loss, dist_loss, out, samples_total, step_counter);
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// dtype: 0 = fp32 inputs, 1 = fp16 inputs, 2 = fp32 inputs with fp64 accumulation.
// raw_density != 0: density holds network outputs s (sigma = exp(min(s,12)-1)),
// otherwise density holds sigma itself. rgba is [4][out_stride], premultiplied.
extern "C" int tn_composite_forward(int dtype, const uint32_t* info, uint32_t stride, uint32_t max_samples,
                                    uint32_t out_stride, float min_transmittance, int raw_density,
                                    const uint32_t* ray_steps, const uint32_t* ray_offset, const void* density,
                                    const void* color, const float* dt, float* rgba, hipStream_t stream) {
    TN_CHECK_ARG(info && ray_steps && ray_offset && density && color && dt && rgba && max_samples <= stride);
    TN_CHECK_ARG(max_samples < (1u << 31));  // sample offsets fit an int
    const uint32_t S = 256;
    const uint32_t waves = (max_samples + S - 1) / S + 1;
    const uint32_t blocks = (waves + kCompositeWaves - 1) / kCompositeWaves;
    switch (dtype) {
        case 0:
            composite_forward_kernel<float, float><<<blocks, 64 * kCompositeWaves, 0, stream>>>(
                info, stride, out_stride, S, min_transmittance, raw_density, ray_steps, ray_offset,
                static_cast<const float*>(density), static_cast<const float*>(color), dt, rgba);
            break;
        case 1:
            composite_forward_kernel<_Float16, float><<<blocks, 64 * kCompositeWaves, 0, stream>>>(
                info, stride, out_stride, S, min_transmittance, raw_density, ray_steps, ray_offset,
                static_cast<const _Float16*>(density), static_cast<const _Float16*>(color), dt, rgba);
            break;
        case 2:
            composite_forward_kernel<float, double><<<blocks, 64 * kCompositeWaves, 0, stream>>>(
                info, stride, out_stride, S, min_transmittance, raw_density, ray_steps, ray_offset,
                static_cast<const float*>(density), static_cast<const float*>(color), dt, rgba);
            break;
        default:
            return TN_ERR_INVALID_ARGUMENT;
    }
    TN_CHECK_LAUNCH();
    return TN_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant for callers that cannot allocate (code under graph
// capture, serving loops): the same kernels with their scratch in a static
// device array instead of a caller-passed workspace. Each GPU the library
// runs on holds its own copy until the process ends; calls that overlap on
// one GPU share it and race.
// ---------------------------------------------------------------------------
namespace tn {
constexpr size_t kStaticScratchBytes = (size_t)32 << 20;
__device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
}  // namespace tn

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch` (in host code the
// array's name is only a stand-in): looked up once per device, with that
// device current, and kept.
int static_scratch(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return TN_ERR_LAUNCH;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(tn::scratch)) != hipSuccess || p == nullptr) return TN_ERR_LAUNCH;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return TN_OK;
}
}  // namespace

extern "C" size_t tn_composite_train_noalloc_max_bytes(void) { return kStaticScratchBytes; }

extern "C" int tn_composite_train_noalloc(int dtype, const uint32_t* info, uint32_t n_rays_max, uint32_t stride,
                                          uint32_t sample_stride, float grad_scale, int scale_by_rays,
                                          float min_transmittance, const uint32_t* ray_steps,
                                          const uint32_t* ray_offset, const float* bg_rgb, const float* target_rgba,
                                          const void* density, const void* color, const float* dt, float* loss,
                                          float* dL_ddensity, void* dL_dcolor, hipStream_t stream) {
    if (tn_composite_scratch_bytes(n_rays_max, sample_stride) > kStaticScratchBytes) return TN_ERR_OUT_OF_MEMORY;
    void* p = nullptr;
    const int rc = static_scratch(&p);
    if (rc != TN_OK) return rc;
    return tn_composite_train(dtype, info, n_rays_max, stride, sample_stride, grad_scale, scale_by_rays,
                              min_transmittance, ray_steps, ray_offset, bg_rgb, target_rgba, density, color, dt, loss,
                              dL_ddensity, dL_dcolor, p, kStaticScratchBytes, stream);
}
