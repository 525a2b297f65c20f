// Core engines: OccupancyGrid, NerfNetwork, Trainer (see core.hpp).
#include "core.hpp"

#include <algorithm>
#include <cstring>
#include <random>

namespace turbo {

// ------------------------------------------------------------------ timing
PhaseTimer::~PhaseTimer() {
    for (auto& r : pending_) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t PhaseTimer::get_() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    TN_HIP(hipEventCreate(&e));
    return e;
}

void PhaseTimer::begin(const char* name, hipStream_t s) {
    Rec r{name, get_(), get_()};
    TN_HIP(hipEventRecord(r.a, s));
    pending_.push_back(r);
}

void PhaseTimer::end(hipStream_t s) { TN_HIP(hipEventRecord(pending_.back().b, s)); }

void PhaseTimer::resolve() {
    for (auto& r : pending_) {
        TN_HIP(hipEventSynchronize(r.b));
        float ms = 0.0f;
        TN_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto& t = totals[r.name];
        t.first += ms;
        t.second += 1;
        pool_.push_back(r.a);
        pool_.push_back(r.b);
    }
    pending_.clear();
}

// ------------------------------------------------------------------ grid
uint32_t OccupancyGrid::count_occupied(hipStream_t s) {
    const size_t nb = tn_popcount_workspace_bytes(volume());
    workspace.reserve(std::max<size_t>(nb, 4));
    count.reserve(1);
    TN_TRY(tn_popcount(bitfield.data(), volume(), count.data(), workspace.data(), workspace.size(), as_tn(s)));
    uint32_t h = 0;
    TN_HIP(hipMemcpyAsync(&h, count.data(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    return h;
}

// ------------------------------------------------------------------ network
void NetworkWorkspace::enlarge(uint32_t B, uint32_t R, int prec) {
    if (B <= batch_size && R <= ray_capacity && prec == precision) return;
    batch_size = std::max(B, batch_size);
    ray_capacity = std::max(R, ray_capacity);
    precision = prec;
    const size_t e = prec == 1 ? 2 : 4;
    // features and their gradients are fp32 in both precisions
    features.resize(32 * (size_t)batch_size * 4);
    density.resize((size_t)batch_size * e);
    rgb.resize(3 * (size_t)batch_size * e);
    dL_drgb.resize(3 * (size_t)batch_size * e);
    dfeat.resize(32 * (size_t)batch_size * 4);
    dL_ds.resize(batch_size);
    dL_dist.resize(batch_size);
    ray_loss.resize(4 * (size_t)ray_capacity);
    ray_dist_loss.resize(ray_capacity);
    loss_out.resize(2);
    info_tail.resize(4);
    composite_scratch.resize(tn_composite_scratch_bytes(ray_capacity, batch_size));
    distortion_scratch.resize(tn_distortion_scratch_bytes(ray_capacity, batch_size));
    mlp_ws.resize(std::max<size_t>(tn_mlp_backward_workspace_bytes(batch_size), 4));
    if (app_dz.size()) app_dz.resize(64 * (size_t)batch_size);
}

size_t NetworkWorkspace::bytes() const {
    return features.bytes() + density.bytes() + rgb.bytes() + dL_drgb.bytes() + dfeat.bytes() + dL_ds.bytes() +
           dL_dist.bytes() + ray_loss.bytes() + ray_dist_loss.bytes() + composite_scratch.bytes() +
           distortion_scratch.bytes() + mlp_ws.bytes() + hash_ws.bytes() + app_dz.bytes() + app_ws.bytes();
}

void NerfNetwork::update_aabb_scale_if_needed(int aabb) {
    if (aabb == aabb_scale) return;
    const double n_min = config.base_resolution;
    const double n_max = (double)config.finest_resolution_per_unit * (double)aabb;
    const double b = std::exp((std::log(n_max) - std::log(n_min)) / (config.n_levels - 1.0));
    TN_TRY(tn_hashgrid_layout(config.n_levels, config.n_features, config.log2_hashmap_size, config.base_resolution, b,
                              &grid));
    aabb_scale = aabb;
}

void NerfNetwork::update_params_if_needed(hipStream_t s, NetworkParamsWorkspace& p) {
    const uint32_t n_net = TN_DENSITY_MLP_PARAMS + TN_COLOR_MLP_PARAMS;
    const uint32_t n_app = n_appearances;
    const uint32_t total = grid.n_params + n_net + (n_app ? TN_APPEARANCE_WEIGHTS + n_app * TN_APPEARANCE_DIMS : 0);
    const bool want_lp = config.precision == 1;
    if (p.n_total == total && p.n_appearances == n_app && p.params_lp.size() == (want_lp ? (size_t)total * 2 : 0))
        return;
    p.n_grid = grid.n_params;
    p.n_total = total;
    p.n_appearances = n_app;
    // seeded initialisation: hash entries uniform in +-hash_init_scale, MLP
    // weights Xavier-uniform per matrix (the colour MLP's first layer over its
    // 48 inputs with an embedding), embeddings uniform in [0, 1) as the
    // reference initialises them (src/core/embedding.cuh:155-157)
    std::vector<float> h(total);
    std::mt19937_64 rng(config.seed);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    for (uint32_t i = 0; i < grid.n_params; ++i) h[i] = config.hash_init_scale * u(rng);
    const int w3_in = n_app ? 32 + TN_APPEARANCE_DIMS : 32;
    const int shapes[5][3] = {{64, 32, 32}, {16, 64, 64}, {64, 32, w3_in}, {64, 64, 64}, {3, 64, 64}};
    uint32_t off = grid.n_params;
    for (auto& sh : shapes) {
        const float a = std::sqrt(6.0f / (float)(sh[0] + sh[2]));
        for (int k = 0; k < sh[0] * sh[1]; ++k) h[off + k] = a * u(rng);
        off += sh[0] * sh[1];
    }
    if (n_app) {
        const float a = std::sqrt(6.0f / (float)(64 + w3_in));
        for (int k = 0; k < TN_APPEARANCE_WEIGHTS; ++k) h[off + k] = a * u(rng);
        off += TN_APPEARANCE_WEIGHTS;
        for (uint32_t k = 0; k < n_app * TN_APPEARANCE_DIMS; ++k) h[off + k] = 0.5f * (u(rng) + 1.0f);
        app_bias.resize((size_t)n_app * 64);
    } else {
        app_bias.release();
    }
    p.params_fp.resize(total);
    TN_HIP(hipMemcpyAsync(p.params_fp.data(), h.data(), total * sizeof(float), hipMemcpyHostToDevice, s));
    if (want_lp) {
        p.params_lp.resize((size_t)total * 2);
        TN_TRY(tn_cast_params(total, p.params_fp.data(), 1, p.params_lp.data(), as_tn(s)));
    } else {
        p.params_lp.release();
    }
    p.gradients.resize(total);
    p.gradients.zero(s);
    p.m1.resize(total);
    p.m1.zero(s);
    p.m2.resize(total);
    p.m2.zero(s);
    p.steps.resize(total);
    p.steps.zero(s);
    TN_HIP(hipStreamSynchronize(s));  // the host staging vector goes out of scope
}

void NerfNetwork::train(hipStream_t s, NetworkParamsWorkspace& p, uint32_t step, uint32_t B, uint32_t R,
                        const uint32_t* info, const float* random_rgb, const uint32_t* ray_steps,
                        const uint32_t* ray_offset, const uint32_t* appearance_id_batch, const float* t_batch,
                        const float* pos_batch, const float* dir_batch, const float* dt_batch,
                        const float* m_norm_batch, const float* dt_norm_batch, const float* target_rgba,
                        const Settings& st) {
    const int prec = config.precision;
    prepare_workspace(B, R);
    NetworkWorkspace& w = workspace;
    const void* params = params_for_kernels(p);
    const tn_stream_t ts = as_tn(s);
    PhaseTimer* tm = st.timer;
    // appearance embedding: this step's per-image bias rows of the colour MLP
    const uint32_t n_app = p.n_appearances;
    tn_appearance_t app_fwd{}, app_bwd{};
    if (n_app) {
        if (!appearance_id_batch) throw std::runtime_error("NerfNetwork::train: the embedding needs image ids");
        ScopedPhase ph(tm, "appearance", s);
        TN_TRY(tn_appearance_bias(n_app, p.params_fp.data() + p.w3e_offset(), p.params_fp.data() + p.emb_offset(),
                                  app_bias.data(), ts));
        app_fwd = tn_appearance_t{app_bias.data(), appearance_id_batch, n_app, nullptr};
        app_bwd = tn_appearance_t{app_bias.data(), appearance_id_batch, n_app, w.app_dz.data()};
    }
    // The per-sample passes run over samples [0, h) on s and, when the batch
    // is split, over [h, B) on the auxiliary stream: the second half's
    // encoding overlaps the first half's MLP forward, and the second half's
    // MLP backward overlaps the first half's hash-grid scatter (MFMA-bound
    // and memory-bound kernels side by side). h is a whole number of the
    // hash-grid backward's count blocks, so each half scatters its own
    // records into the batch's plan; compositing, the sums and Adam run once.
    int halves_knob = 0;
    TN_TRY(tn_tune_get(TN_TUNE_STEP_HALVES, &halves_knob));
    const uint32_t part = st.aux && st.hash_plan && halves_knob ? tn_hashgrid_backward_part_samples(B, grid) : 0u;
    const uint32_t h = part ? ((B + part - 1) / part / 2) * part : 0u;
    const bool split = h > 0 && h < B;
    const size_t esz = prec == 1 ? 2 : 4;
    auto at = [](auto* p, size_t bytes) { return reinterpret_cast<decltype(p)>(reinterpret_cast<uintptr_t>(p) + bytes); };
    // half k: its first sample, sample bound, descriptor and stream
    struct Half { uint32_t first, n; const uint32_t* info; hipStream_t s; const char* tag; };
    const Half halves[2] = {{0, split ? h : B, info, s, ""}, {h, B - h, w.info_tail.data(), st.aux, "/2"}};
    const int n_halves = split ? 2 : 1;
    // the second half's phases are timed under their own names ("…/2")
    auto phase = [](const Half& hf, const char* name) { return std::string(name) + hf.tag; };
    if (split) {
        TN_TRY(tn_batch_tail(info, h, w.info_tail.data(), ts));
        TN_HIP(hipEventRecord(st.split_events[0], s));
        TN_HIP(hipStreamWaitEvent(st.aux, st.split_events[0], 0));
    }
    // forward: encoding, density + colour MLPs (reference :431-545)
    for (int k = 0; k < n_halves; ++k) {
        const Half& hf = halves[k];
        const tn_stream_t hs = as_tn(hf.s);
        tn_appearance_t af = app_fwd;
        if (n_app) af.image_id = appearance_id_batch + hf.first;
        {
            ScopedPhase ph(tm, phase(hf, "hashgrid_forward").c_str(), hf.s);
            TN_TRY(tn_hashgrid_forward(prec, 0, hf.n, hf.info, B, B, grid, pos_batch + hf.first, params,
                                       at(w.features.data(), 4 * (size_t)hf.first), hs));
        }
        {
            ScopedPhase ph(tm, phase(hf, "mlp_forward").c_str(), hf.s);
            TN_TRY(tn_mlp_forward(prec, hf.n, hf.info, B, B, B, params, p.n_grid,
                                  at(w.features.data(), 4 * (size_t)hf.first), dir_batch + hf.first, 1,
                                  at(w.density.data(), esz * hf.first), at(w.rgb.data(), esz * hf.first),
                                  n_app ? &af : nullptr, hs));
        }
    }
    if (split) {
        TN_HIP(hipEventRecord(st.split_events[1], st.aux));
        TN_HIP(hipStreamWaitEvent(s, st.split_events[1], 0));
    }
    // distortion loss (reference :273-283)
    if (st.use_distortion_loss) {
        ScopedPhase ph(tm, "distortion", s);
        TN_TRY(tn_distortion(prec, info, R, B, st.distortion_lambda, ray_steps, ray_offset, w.density.data(),
                             m_norm_batch, dt_norm_batch, w.ray_dist_loss.data(), w.dL_dist.data(),
                             w.distortion_scratch.data(), w.distortion_scratch.size(), ts));
    }
    // compositing + loss + its backward (reference :547-603)
    {
        ScopedPhase ph(tm, "composite", s);
        // gradients scaled by loss_scale * 2^ceil(log2(rays)), unscaled by Adam
        TN_TRY(tn_composite_train(prec, info, R, R, B, config.loss_scale(), 1, st.min_transmittance, ray_steps,
                                  ray_offset, random_rgb, target_rgba, w.density.data(), w.rgb.data(), dt_batch,
                                  w.ray_loss.data(), w.dL_ds.data(), w.dL_drgb.data(), w.composite_scratch.data(),
                                  w.composite_scratch.size(), ts));
    }
    // network backward (reference :668-768) and the hash-grid records of
    // each half; the second half's weight gradients add to the first's
    const size_t hb = tn_hashgrid_backward_workspace_bytes(B, grid);
    w.hash_ws.reserve(hb);
    for (int k = n_halves - 1; k >= 0; --k) {
        // the second half's MLP backward waits for the first's (one weight
        // gradient, one partial-sum workspace), then scatters its records
        const Half& hf = halves[split ? 1 - k : 0];
        const tn_stream_t hs = as_tn(hf.s);
        tn_appearance_t ab = app_bwd;
        if (n_app) {
            ab.image_id = appearance_id_batch + hf.first;
            ab.dz = w.app_dz.data() + hf.first;
        }
        if (split && hf.first) TN_HIP(hipStreamWaitEvent(hf.s, st.split_events[2], 0));
        {
            ScopedPhase ph(tm, phase(hf, "mlp_backward").c_str(), hf.s);
            TN_TRY(tn_mlp_backward(prec, hf.n, hf.info, B, B, params, p.n_grid,
                                   at(w.features.data(), 4 * (size_t)hf.first), dir_batch + hf.first,
                                   t_batch + hf.first, at(w.dL_drgb.data(), esz * hf.first), B,
                                   w.dL_ds.data() + hf.first,
                                   st.use_distortion_loss ? w.dL_dist.data() + hf.first : nullptr, config.loss_scale(),
                                   1, at(w.dfeat.data(), 4 * (size_t)hf.first), p.gradients.data() + p.n_grid,
                                   hf.first ? 1 : 0, w.mlp_ws.data(), w.mlp_ws.size(), n_app ? &ab : nullptr, hs));
        }
        if (split && !hf.first) TN_HIP(hipEventRecord(st.split_events[2], s));
        if (split) {
            // the records of this half's samples into the batch's plan
            ScopedPhase ph(tm, phase(hf, "hashgrid_scatter").c_str(), hf.s);
            TN_TRY(tn_hashgrid_backward_scatter(0, B, info, B, B, grid, pos_batch, w.dfeat.data(), st.hash_plan,
                                                st.hash_plan_bytes, w.hash_ws.data(), w.hash_ws.size(), hf.first,
                                                hf.first ? B : h, hs));
        }
    }
    if (split) {
        TN_HIP(hipEventRecord(st.split_events[3], st.aux));
        TN_HIP(hipStreamWaitEvent(s, st.split_events[3], 0));
    }
    if (n_app) {
        ScopedPhase ph(tm, "appearance", s);
        TN_TRY(tn_appearance_backward(B, info, B, appearance_id_batch, w.app_dz.data(), n_app,
                                      p.params_fp.data() + p.w3e_offset(), p.params_fp.data() + p.emb_offset(),
                                      p.gradients.data() + p.w3e_offset(), p.gradients.data() + p.emb_offset(),
                                      w.app_ws.data(), w.app_ws.bytes(), ts));
    }
    // optimiser settings (reference :770-781, lr schedule :256; from the
    // device step counter in a captured step)
    const bool dev_lr = st.step_for_lr && st.step_counter;
    const float lr = dev_lr ? config.learning_rate
                            : config.learning_rate / (1.0f + config.learning_rate_decay * (float)step);
    // With the batch's plan, the hash table takes its Adam step inside the
    // hash-grid backward's sum passes (its gradient is never written and
    // re-read); the separate Adam pass then covers the network weights and
    // embeddings only. Same parameters, moments and steps bit for bit.
    int fused_knob = 0;
    TN_TRY(tn_tune_get(TN_TUNE_HASH_ADAM_FUSED, &fused_knob));
    const bool fused = fused_knob && st.hash_plan;
    const tn_adam_t adam_hash{config.loss_scale(), info, lr, dev_lr ? st.step_counter : nullptr,
                              config.learning_rate_decay, config.beta1, config.beta2, config.epsilon,
                              p.params_fp.data(), prec, prec == 1 ? p.params_lp.data() : nullptr, p.m1.data(),
                              p.m2.data(), p.steps.data()};
    {
        ScopedPhase ph(tm, "hashgrid_backward", s);
        // the bin-sort scatter (its workspace follows the batch size and the
        // layout); with the batch's plan formed ahead, only the passes that
        // need the gradients run here
        if (split && fused)
            TN_TRY(tn_hashgrid_backward_accumulate_adam(B, grid, p.gradients.data(), st.hash_plan,
                                                        st.hash_plan_bytes, w.hash_ws.data(), w.hash_ws.size(),
                                                        &adam_hash, ts));
        else if (split)
            TN_TRY(tn_hashgrid_backward_accumulate(B, grid, p.gradients.data(), st.hash_plan, st.hash_plan_bytes,
                                                   w.hash_ws.data(), w.hash_ws.size(), ts));
        else if (fused)
            TN_TRY(tn_hashgrid_backward_planned_adam(0, B, info, B, B, grid, pos_batch, w.dfeat.data(),
                                                     p.gradients.data(), st.hash_plan, st.hash_plan_bytes,
                                                     w.hash_ws.data(), w.hash_ws.size(), &adam_hash, ts));
        else if (st.hash_plan)
            TN_TRY(tn_hashgrid_backward_planned(0, B, info, B, B, grid, pos_batch, w.dfeat.data(), p.gradients.data(),
                                                st.hash_plan, st.hash_plan_bytes, w.hash_ws.data(), w.hash_ws.size(),
                                                ts));
        else
            TN_TRY(tn_hashgrid_backward_binned(0, B, info, B, B, grid, pos_batch, w.dfeat.data(), p.gradients.data(),
                                               w.hash_ws.data(), w.hash_ws.size(), ts));
    }
    {
        ScopedPhase ph(tm, "adam", s);
        // the parameters from `first` on: everything, or past the hash table
        // when its step is fused above
        const uint32_t first = fused ? p.n_grid : 0u;
        uint8_t* lp = prec == 1 ? p.params_lp.data() + 2 * (size_t)first : nullptr;
        TN_TRY(tn_adam_step(p.n_total - first, p.n_grid - first, p.n_net_end() - first, config.loss_scale(), info, lr,
                            dev_lr ? st.step_counter : nullptr, config.learning_rate_decay, config.beta1,
                            config.beta2, config.epsilon, config.l2_reg, p.params_fp.data() + first, prec, lp,
                            p.gradients.data() + first, p.m1.data() + first, p.m2.data() + first,
                            p.steps.data() + first, ts));
    }
    // losses: mean over rays of the four smooth-L1 channels, mean distortion,
    // and the running count of trained samples
    {
        ScopedPhase ph(tm, "loss_reduce", s);
        TN_TRY(tn_step_metrics(info, R, w.ray_loss.data(), st.use_distortion_loss ? w.ray_dist_loss.data() : nullptr,
                               w.loss_out.data(), st.samples_total, st.step_counter, ts));
    }
}

void NerfNetwork::prepare_workspace(uint32_t B, uint32_t R) {
    workspace.enlarge(B, R, config.precision);
    workspace.hash_ws.reserve(tn_hashgrid_backward_workspace_bytes(B, grid));
    if (n_appearances) {
        if (workspace.app_dz.size() < 64 * (size_t)workspace.batch_size) workspace.app_dz.resize(64 * (size_t)workspace.batch_size);
        const size_t ab = tn_appearance_workspace_bytes(n_appearances) / sizeof(float);
        if (workspace.app_ws.size() < ab) workspace.app_ws.resize(ab);
    }
}

void NerfNetwork::inference(hipStream_t s, const NetworkParamsWorkspace& p, uint32_t n, const uint32_t* info,
                            uint32_t stride, const float* pos_batch, const float* dir_batch, void* features,
                            void* density_out, void* rgb_out, bool use_color_network) {
    const int prec = config.precision;
    const void* params = params_for_kernels(p);
    const int fdt = prec == 1 && config.inference_fp16_features ? 1 : 0;  // feature dtype
    TN_TRY(tn_hashgrid_forward(prec, fdt, n, info, stride, stride, grid, pos_batch, params, features, as_tn(s)));
    tn_appearance_t app{};
    if (use_color_network && p.n_appearances) {
        TN_TRY(tn_appearance_bias(1, p.params_fp.data() + p.w3e_offset(), p.params_fp.data() + p.emb_offset(),
                                  app_bias.data(), as_tn(s)));
        app = tn_appearance_t{app_bias.data(), nullptr, 1, nullptr};
    }
    TN_TRY(tn_mlp_inference(prec, fdt, n, info, stride, stride, stride, params, p.n_grid, features, dir_batch,
                            use_color_network ? 1 : 0, density_out, rgb_out, app.bias ? &app : nullptr, as_tn(s)));
}

// ------------------------------------------------------------------ workspaces
void TrainingWorkspace::enlarge(uint32_t B, uint32_t R) {
    if (B == batch_size && R == ray_capacity) return;
    batch_size = B;
    ray_capacity = R;
    random_float.resize(4 * (size_t)R);
    pix_rgba.resize(4 * (size_t)R);
    ray_origin.resize(3 * (size_t)R);
    ray_dir.resize(3 * (size_t)R);
    ray_t.resize(R);
    ray_t_max.resize(R);
    ray_img_id.resize(R);
    ray_step.resize(R);
    ray_offset.resize(R);
    ray_alive.resize(R);
    sample_img_id.resize(B);
    sample_t.resize(B);
    sample_pos.resize(3 * (size_t)B);
    sample_dir.resize(3 * (size_t)B);
    sample_dt.resize(B);
    sample_m_norm.resize(B);
    sample_dt_norm.resize(B);
    info.resize(4);
    scan_ws.resize(std::max<size_t>(tn_scan_workspace_bytes(R), 4));
}

size_t TrainingWorkspace::bytes() const {
    return random_float.bytes() + pix_rgba.bytes() + ray_origin.bytes() + ray_dir.bytes() + ray_t.bytes() +
           ray_t_max.bytes() + ray_img_id.bytes() + ray_step.bytes() + ray_offset.bytes() + ray_alive.bytes() +
           sample_img_id.bytes() + sample_t.bytes() + sample_pos.bytes() + sample_dir.bytes() + sample_dt.bytes() +
           sample_m_norm.bytes() + sample_dt_norm.bytes() + scan_ws.bytes() + hash_plan.bytes();
}

void TrainingWorkspace::free_allocations() { *this = TrainingWorkspace(); }

void DatasetWorkspace::upload(const Dataset& ds, hipStream_t s) {
    n_images = (uint32_t)ds.n_images();
    image_dims = ds.image_dimensions;
    n_pixels_per_image = ds.n_pixels_per_image;
    std::vector<tn_camera_t> cams(n_images);
    for (uint32_t i = 0; i < n_images; ++i) cams[i] = ds.cameras[i].c();
    cameras.resize(n_images * sizeof(tn_camera_t));
    TN_HIP(hipMemcpyAsync(cameras.data(), cams.data(), cameras.size(), hipMemcpyHostToDevice, s));
    images.resize(ds.image_data.size());
    TN_HIP(hipMemcpyAsync(images.data(), ds.image_data.data(), images.size(), hipMemcpyHostToDevice, s));
    const tn_bbox_t b = ds.bounding_box.c();
    bbox.resize(sizeof(b));
    TN_HIP(hipMemcpyAsync(bbox.data(), &b, sizeof(b), hipMemcpyHostToDevice, s));
    TN_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ trainer
Trainer::Context::Context(const Dataset* ds, NeRF* n, uint32_t B) : dataset(ds), nerf(n), batch_size(B) {
    TN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    TN_HIP(hipStreamCreateWithFlags(&gen_stream, hipStreamNonBlocking));
    TN_HIP(hipStreamCreateWithFlags(&aux_stream, hipStreamNonBlocking));
    for (auto& e : ev_split) TN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (int k = 0; k < 2; ++k) {
        TN_HIP(hipEventCreateWithFlags(&ev_ready[k], hipEventDisableTiming));
        TN_HIP(hipEventCreateWithFlags(&ev_consumed[k], hipEventDisableTiming));
    }
    TN_HIP(hipEventCreateWithFlags(&ev_grid, hipEventDisableTiming));
    TN_HIP(hipEventCreateWithFlags(&ev_gen_idle, hipEventDisableTiming));
}

void Trainer::Context::destroy_graphs() {
    for (int c = 0; c < 2; ++c) {
        if (train_graph[c]) (void)hipGraphExecDestroy(train_graph[c]);
        if (gen_graph[c]) (void)hipGraphExecDestroy(gen_graph[c]);
        train_graph[c] = gen_graph[c] = nullptr;
    }
    graph_key.clear();
}

Trainer::Context::~Context() {
    (void)hipStreamSynchronize(gen_stream);
    (void)hipStreamSynchronize(stream);
    (void)hipStreamSynchronize(aux_stream);
    for (int k = 0; k < 2; ++k) {
        (void)hipEventDestroy(ev_ready[k]);
        (void)hipEventDestroy(ev_consumed[k]);
    }
    for (auto e : ev_split) (void)hipEventDestroy(e);
    destroy_graphs();
    (void)hipEventDestroy(ev_grid);
    (void)hipEventDestroy(ev_gen_idle);
    (void)hipStreamDestroy(gen_stream);
    (void)hipStreamDestroy(aux_stream);
    (void)hipStreamDestroy(stream);
}

void Trainer::generate_next_training_batch(Context& ctx, TrainingWorkspace& ws, hipStream_t s, PhaseTimer* tm) {
    NeRF& nerf = *ctx.nerf;
    const tn_stream_t ts = as_tn(s);
    const uint32_t R = ws.ray_capacity, B = ws.batch_size;
    const Dataset& ds = *ctx.dataset;
    const DatasetWorkspace& dw = nerf.dataset_ws;
    {
        // seeded uniforms: slice 0 picks pixels, slices 1-3 are the background
        // colour (reference trainer.cu:45-50,332); rays are generated for the
        // device-side estimate of how many fill the batch and split evenly
        // over the images by that count (-1 below)
        ScopedPhase ph(tm, "rays", s);
        TN_TRY(tn_fill_uniform_dev(4 * (uint64_t)R, ctx.rng_seed, ctx.d_rng.data(), ws.random_float.data(), ts));
        TN_TRY(tn_init_training_rays(R, ctx.n_rays_next.data(), R, dw.n_images, dw.n_pixels_per_image,
                                     dw.image_dims.x, dw.image_dims.y, -1.0f,
                                     reinterpret_cast<const tn_bbox_t*>(dw.bbox.data()),
                                     reinterpret_cast<const tn_camera_t*>(dw.cameras.data()), dw.images.data(),
                                     ws.random_float.data(), ws.pix_rgba.data(), ws.ray_img_id.data(),
                                     ws.ray_origin.data(), ws.ray_dir.data(), ws.ray_t.data(), ws.ray_t_max.data(),
                                     ws.ray_alive.data(), ts));
        if (ctx.alpha_selection_threshold < 1.0f && ctx.alpha_selection_probability < 1.0f)
            TN_TRY(tn_deactivate_rays(R, ctx.n_rays_next.data(), R, ctx.alpha_selection_threshold,
                                      ctx.alpha_selection_probability, ws.pix_rgba.data(), ws.random_float.data() + R,
                                      ws.ray_alive.data(), ts));
    }
    const float dt_min = ctx.min_step_size;
    const float dt_max = ds.bounding_box.size() * dt_min;
    const tn_bbox_t bbox = ds.bounding_box.c();
    const tn_occupancy_grid_t grid = nerf.occupancy_grid.view();
    {
        ScopedPhase ph(tm, "march_count", s);
        TN_TRY(tn_march_count(R, ctx.n_rays_next.data(), R, bbox, grid, NeRFConstants::cone_angle, dt_min, dt_max,
                              0xFFFFFFFFu, ws.ray_dir.data(), ws.ray_t_max.data(), ws.ray_alive.data(),
                              ws.ray_origin.data(), ws.ray_t.data(), ws.ray_step.data(), ts));
    }
    {
        ScopedPhase ph(tm, "scan_select", s);
        TN_TRY(tn_scan_select(0, R, ctx.n_rays_next.data(), ws.ray_step.data(), B, ws.ray_offset.data(),
                              ws.info.data(), ws.scan_ws.data(), ws.scan_ws.size(), ts));
    }
    {
        ScopedPhase ph(tm, "march_generate", s);
        TN_TRY(tn_march_generate(R, ws.info.data(), R, B, bbox, 1.0f / ds.bounding_box.size(), grid, dt_min, dt_max,
                                 NeRFConstants::cone_angle, ws.ray_img_id.data(), ws.ray_origin.data(),
                                 ws.ray_dir.data(), ws.ray_t.data(), ws.ray_offset.data(), ws.ray_alive.data(),
                                 ws.ray_step.data(), ws.sample_img_id.data(), ws.sample_t.data(), ws.sample_pos.data(),
                                 ws.sample_dir.data(), ws.sample_dt.data(), ws.sample_m_norm.data(),
                                 ws.sample_dt_norm.data(), ts));
        TN_TRY(tn_estimate_rays(ws.info.data(), B, R, 64, ctx.n_rays_next.data(), ts));
    }
    {
        // the hash-grid backward's plan needs only the positions: formed here,
        // beside the network stream's work, rather than on it
        ScopedPhase ph(tm, "hashgrid_plan", s);
        const tn_hashgrid_t& g = nerf.network.grid;
        const size_t hp = tn_hashgrid_backward_plan_bytes(B, g);
        ws.hash_plan.reserve(hp);  // sized ahead of any capture (prepare_plans)
        if (hp)
            TN_TRY(tn_hashgrid_backward_plan(B, ws.info.data(), B, g, ws.sample_pos.data(), ws.hash_plan.data(),
                                             ws.hash_plan.size(), ts));
    }
}

void Trainer::prepare_plans(Context& ctx) {
    for (auto& b : ctx.batches) b.hash_plan.reserve(tn_hashgrid_backward_plan_bytes(b.batch_size, ctx.nerf->network.grid));
}

// Train on batches[c] on the network stream; with `graph` the work is being
// captured.
void Trainer::train_network(Context& ctx, int c, uint32_t training_step, bool graph) {
    TrainingWorkspace& ws = ctx.batches[c];
    NerfNetwork::Settings st;
    st.use_distortion_loss = ctx.use_distortion_loss;
    st.distortion_lambda = ctx.distortion_lambda;
    st.timer = graph ? nullptr : &ctx.timer;
    st.samples_total = ctx.samples_total.size() ? ctx.samples_total.data() : nullptr;
    st.step_counter = ctx.d_step.data();
    st.step_for_lr = graph;
    if (ws.hash_plan.size() && ws.hash_plan.size() >= tn_hashgrid_backward_plan_bytes(ws.batch_size, ctx.nerf->network.grid)) {
        st.hash_plan = ws.hash_plan.data();
        st.hash_plan_bytes = ws.hash_plan.size();
        st.aux = ctx.aux_stream;
        st.split_events = ctx.ev_split;
    }
    ctx.nerf->network.train(ctx.stream, ctx.nerf->params, training_step, ws.batch_size, ws.ray_capacity,
                            ws.info.data(), ws.random_float.data() + ws.ray_capacity, ws.ray_step.data(),
                            ws.ray_offset.data(), ws.sample_img_id.data(), ws.sample_t.data(), ws.sample_pos.data(),
                            ws.sample_dir.data(), ws.sample_dt.data(), ws.sample_m_norm.data(),
                            ws.sample_dt_norm.data(), ws.pix_rgba.data(), st);
}

// Eager step: batches[c ^ 1] is generated on the generation stream once its
// last training has consumed it, while batches[c] trains on the network
// stream once it is ready.
void Trainer::train_batch(Context& ctx, int c, uint32_t training_step) {
    const int nx = c ^ 1;
    TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ctx.ev_consumed[nx], 0));
    generate_next_training_batch(ctx, ctx.batches[nx], ctx.gen_stream, &ctx.timer);
    TN_HIP(hipEventRecord(ctx.ev_ready[nx], ctx.gen_stream));
    TN_HIP(hipStreamWaitEvent(ctx.stream, ctx.ev_ready[c], 0));
    train_network(ctx, c, training_step, false);
}

void Trainer::enqueue_train_step(Context& ctx, uint32_t training_step) {
    const int c = ctx.cur;
    if (!ctx.prefetched) {
        // first step, or the grid changed since the batch generated ahead
        TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ctx.ev_grid, 0));
        TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ctx.ev_consumed[c], 0));
        generate_next_training_batch(ctx, ctx.batches[c], ctx.gen_stream, &ctx.timer);
        TN_HIP(hipEventRecord(ctx.ev_ready[c], ctx.gen_stream));
    }
    train_batch(ctx, c, training_step);
    TN_HIP(hipEventRecord(ctx.ev_consumed[c], ctx.stream));
    ctx.last = c;
    ctx.cur = c ^ 1;
    ctx.prefetched = true;
}

// What a captured step has baked into its kernel arguments: buffer addresses,
// sizes and settings. A change of any of them recaptures.
std::vector<uint64_t> Trainer::graph_key(const Context& ctx) const {
    const NeRF& n = *ctx.nerf;
    const NetworkConfig& c = n.network.config;
    auto bits = [](float f) { uint32_t u; std::memcpy(&u, &f, 4); return (uint64_t)u; };
    auto ptr = [](const void* p) { return (uint64_t)reinterpret_cast<uintptr_t>(p); };
    std::vector<uint64_t> k = {
        ptr(ctx.nerf), ptr(ctx.dataset), ctx.batch_size, ctx.batches[0].ray_capacity, ctx.batches[0].batch_size,
        bits(ctx.alpha_selection_threshold), bits(ctx.alpha_selection_probability), bits(ctx.min_step_size),
        (uint64_t)ctx.use_distortion_loss, bits(ctx.distortion_lambda), ctx.rng_seed,
        ptr(n.params.params_fp.data()), ptr(n.params.params_lp.data()), ptr(n.params.gradients.data()),
        ptr(n.params.m1.data()), ptr(n.params.m2.data()), ptr(n.params.steps.data()), n.params.n_total, n.params.n_grid,
        ptr(n.occupancy_grid.bitfield.data()), (uint64_t)n.occupancy_grid.n_levels, ptr(n.dataset_ws.images.data()),
        ptr(n.dataset_ws.cameras.data()), n.dataset_ws.n_images, (uint64_t)c.precision, bits(c.learning_rate),
        bits(c.learning_rate_decay), bits(c.beta1), bits(c.beta2), bits(c.epsilon), bits(c.l2_reg),
        (uint64_t)n.network.aabb_scale, n.network.grid.n_params, bits(ctx.dataset->bounding_box.size()),
        ptr(n.network.workspace.features.data()), ptr(n.network.workspace.hash_ws.data()),
        ptr(n.network.workspace.mlp_ws.data()), ptr(ctx.samples_total.data()), tn_tune_generation(),
        n.params.n_appearances, ptr(n.network.app_bias.data()), ptr(n.network.workspace.app_dz.data()),
        ptr(n.network.workspace.app_ws.data())};
    for (const auto& b : ctx.batches) {
        k.push_back(ptr(b.random_float.data()));
        k.push_back(ptr(b.hash_plan.data()));
    }
    return k;
}

namespace {

// `enqueue` captured on `s` (the caller owns the returned graph)
template <typename F>
hipGraph_t capture_on(hipStream_t s, F&& enqueue) {
    hipGraph_t g = nullptr;
    TN_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    TN_HIP(hipStreamEndCapture(s, &g));
    return g;
}

hipGraphExec_t instantiate(hipGraph_t g) {
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TN_HIP(e);
    return x;
}

}  // namespace

hipGraph_t Trainer::capture_train(Context& ctx, int c) {
    return capture_on(ctx.stream, [&] { train_network(ctx, c, 0, true); });
}

hipGraph_t Trainer::capture_gen(Context& ctx, int c) {
    return capture_on(ctx.gen_stream, [&] { generate_next_training_batch(ctx, ctx.batches[c], ctx.gen_stream, nullptr); });
}

void Trainer::instantiate_graphs(Context& ctx, int c) {
    prepare_plans(ctx);
    ctx.train_graph[c] = instantiate(capture_train(ctx, c));
    ctx.gen_graph[c ^ 1] = instantiate(capture_gen(ctx, c ^ 1));
    ++ctx.n_captures;
}

void Trainer::enqueue_train_steps_graph(Context& ctx, uint32_t first_step, uint32_t n) {
    if (n == 0) return;
    begin_graph_steps(ctx, first_step);
    for (uint32_t k = 0; k < n; ++k) enqueue_graph_step(ctx);
}

void Trainer::begin_graph_steps(Context& ctx, uint32_t first_step) {
    const hipStream_t net = ctx.stream;
    // a captured step may not allocate: size every buffer first
    ctx.nerf->network.prepare_workspace(ctx.batches[0].batch_size, ctx.batches[0].ray_capacity);
    prepare_plans(ctx);
    const std::vector<uint64_t> key = graph_key(ctx);
    if (key != ctx.graph_key) {
        ctx.destroy_graphs();
        ctx.graph_key = key;
    }
    // the device step counter takes the host's training step
    TN_HIP(hipStreamSynchronize(net));  // the pinned staging value is free
    ctx.step_host[0] = first_step;
    TN_HIP(hipMemcpyAsync(ctx.d_step.data(), ctx.step_host.data(), sizeof(uint32_t), hipMemcpyHostToDevice, net));
}

void Trainer::enqueue_graph_step(Context& ctx) {
    const hipStream_t net = ctx.stream, gen = ctx.gen_stream;
    const int c = ctx.cur, nx = c ^ 1;
    if (!ctx.prefetched) {
        TN_HIP(hipStreamWaitEvent(gen, ctx.ev_grid, 0));
        TN_HIP(hipStreamWaitEvent(gen, ctx.ev_consumed[c], 0));
        generate_next_training_batch(ctx, ctx.batches[c], gen, &ctx.timer);
        TN_HIP(hipEventRecord(ctx.ev_ready[c], gen));
    }
    if (!ctx.train_graph[c]) instantiate_graphs(ctx, c);
    // the next batch, once its last training has consumed it
    TN_HIP(hipStreamWaitEvent(gen, ctx.ev_consumed[nx], 0));
    TN_HIP(hipGraphLaunch(ctx.gen_graph[nx], gen));
    TN_HIP(hipEventRecord(ctx.ev_ready[nx], gen));
    // this batch, once it is ready
    TN_HIP(hipStreamWaitEvent(net, ctx.ev_ready[c], 0));
    TN_HIP(hipGraphLaunch(ctx.train_graph[c], net));
    TN_HIP(hipEventRecord(ctx.ev_consumed[c], net));
    ctx.last = c;
    ctx.cur = nx;
    ctx.prefetched = true;
}

void Trainer::read_metrics(Context& ctx, float& loss) {
    const hipStream_t s = ctx.stream;
    TN_HIP(hipMemcpyAsync(ctx.info_host.data(), ctx.batches[ctx.last].info.data(), 4 * sizeof(uint32_t),
                          hipMemcpyDeviceToHost, s));
    TN_HIP(hipMemcpyAsync(ctx.loss_host.data(), ctx.nerf->network.workspace.loss_out.data(), 2 * sizeof(float),
                          hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    ctx.n_rays_in_batch = ctx.info_host[0];
    ctx.n_samples_in_batch = ctx.info_host[1];
    loss = ctx.loss_host[0] + ctx.loss_host[1];
}

float Trainer::train_step(Context& ctx, uint32_t training_step) {
    enqueue_train_step(ctx, training_step);
    float loss = 0.0f;
    read_metrics(ctx, loss);
    if (ctx.info_host[2] != 0) throw std::runtime_error(tn_status_name((int)ctx.info_host[2]));
    return loss;
}

uint32_t Trainer::update_occupancy_grid(Context& ctx, uint32_t training_step) {
    NeRF& nerf = *ctx.nerf;
    OccupancyGrid& og = nerf.occupancy_grid;
    const hipStream_t s = ctx.stream;
    // marching of the batch generated ahead reads the grid: let it finish,
    // then discard that batch (its workspace serves as scratch here) and
    // make the next generation wait for the new grid
    TN_HIP(hipEventRecord(ctx.ev_gen_idle, ctx.gen_stream));
    TN_HIP(hipStreamWaitEvent(s, ctx.ev_gen_idle, 0));
    ctx.prefetched = false;
    TrainingWorkspace& ws = ctx.batches[ctx.cur];
    const tn_stream_t ts = as_tn(s);
    const float aabb = ctx.dataset->bounding_box.size();
    const uint32_t vol = og.volume();
    const uint32_t B = ws.batch_size;
    nerf.network.workspace.enlarge(B, ws.ray_capacity, nerf.network.config.precision);
    NetworkWorkspace& nw = nerf.network.workspace;
    TN_TRY(tn_occ_decay(og.view(), NeRFConstants::occupancy_decay, ts));
    // per cascade, per chunk of at most B cells: jittered points, density
    // inference, max-update (reference trainer.cu:236-295)
    const uint32_t chunk = std::min(B, ws.ray_capacity);
    for (int level = 0; level < og.n_levels; ++level) {
        for (uint32_t start = 0; start < vol; start += chunk) {
            const uint32_t n = std::min(chunk, vol - start);
            // the occupancy update draws from its own stream (key seed ^ k)
            TN_TRY(tn_fill_uniform(4 * (uint64_t)chunk, ctx.rng_seed ^ 0x0cc0cc0cc0cc0cc0ull, ctx.occ_rng_offset,
                                   ws.random_float.data(), ts));
            ctx.occ_rng_offset += 4 * (uint64_t)chunk;
            TN_TRY(tn_occ_sample_points(n, chunk, start, og.view(), level, 1.0f / aabb, ws.random_float.data(),
                                        ws.sample_pos.data(), ts));
            nerf.network.inference(s, nerf.params, n, nullptr, chunk, ws.sample_pos.data(), nullptr,
                                   nw.features.data(), nw.density.data(), nullptr, false);
            TN_TRY(tn_occ_update_density(nerf.network.config.precision, n, start, og.view(), level,
                                         training_step < 256 ? 1 : 0, ws.random_float.data() + 3 * (size_t)chunk,
                                         nw.density.data(), ts));
        }
    }
    TN_TRY(tn_occ_update_bits(og.view(), ctx.occupancy_threshold, ts));
    TN_HIP(hipEventRecord(ctx.ev_grid, s));
    return og.count_occupied(s);
}

}  // namespace turbo
