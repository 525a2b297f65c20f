// NeRFManager and NeRFTrainingController (see controllers.hpp).
#include "controllers.hpp"

#include <algorithm>
#include <new>

namespace turbo {

NeRFProxy* NeRFManager::create() {
    for (size_t i = 0; i < proxies_.size(); ++i) {
        if (!proxies_[i] || !proxies_[i]->is_valid) {
            // a destroyed slot is re-initialised in place: handles to it stay
            // valid (they now refer to the new NeRF) instead of dangling, and
            // the NeRF storage does not move, so no pointer into it dangles
            if (!proxies_[i]) {
                proxies_[i] = std::make_unique<NeRFProxy>();
            } else {
                proxies_[i]->reset();
            }
            proxies_[i]->id = (int)i;
            proxies_[i]->is_valid = true;
            return proxies_[i].get();
        }
    }
    throw std::runtime_error("NeRFManager: no free NeRF slots");
}

NeRFProxy* NeRFManager::clone(NeRFProxy* src) {
    if (!src) throw std::runtime_error("NeRFManager::clone: null NeRF");
    NeRFProxy* p = create();
    // the dataset is not cloned; the trained state is copied (reference nerf-manager.cuh:89-121)
    p->render_bbox = src->render_bbox;
    p->training_bbox = src->training_bbox;
    p->transform = src->transform;
    p->is_visible = src->is_visible;
    p->can_render = src->can_render;
    p->training_step = src->training_step;
    p->clone_source = src->clone_source ? src->clone_source : src;
    const hipStream_t s = DeviceManager::stream();
    NeRF& d = p->nerfs[0];
    const NeRF& o = src->nerfs[0];
    d.network.config = o.network.config;
    d.network.update_aabb_scale_if_needed(o.network.aabb_scale);
    auto copy = [&](auto& dst, const auto& from) {
        dst.resize(from.size());
        if (from.size()) TN_HIP(hipMemcpyAsync(dst.data(), from.data(), from.bytes(), hipMemcpyDeviceToDevice, s));
    };
    d.params.n_grid = o.params.n_grid;
    d.params.n_total = o.params.n_total;
    copy(d.params.params_fp, o.params.params_fp);
    copy(d.params.params_lp, o.params.params_lp);
    d.occupancy_grid.n_levels = o.occupancy_grid.n_levels;
    d.occupancy_grid.resolution = o.occupancy_grid.resolution;
    copy(d.occupancy_grid.bitfield, o.occupancy_grid.bitfield);
    copy(d.occupancy_grid.density, o.occupancy_grid.density);
    TN_HIP(hipStreamSynchronize(s));
    return p;
}

void NeRFManager::destroy(NeRFProxy* p) {
    for (auto& q : proxies_) {
        if (q.get() == p) {
            bool held = false;
            for (auto& n : q->nerfs) held = held || n.holds_device_memory();
            // no queued work on any stream may still use the buffers
            if (held) TN_HIP(hipDeviceSynchronize());
            q->free_device_memory();
            q->is_valid = false;
            ++q->generation;
            return;
        }
    }
}

std::vector<NeRFProxy*> NeRFManager::get_proxies() {
    std::vector<NeRFProxy*> v;
    for (auto& q : proxies_)
        if (q && q->is_valid) v.push_back(q.get());
    return v;
}

void set_analytic_field(NeRFProxy* proxy, const float* blobs_host, uint32_t n_blobs, float aabb, float threshold) {
    if (!proxy || !blobs_host || n_blobs == 0) throw std::runtime_error("set_analytic_field: empty field");
    const hipStream_t s = DeviceManager::stream();
    NeRF& nerf = proxy->nerfs[0];
    nerf.analytic_blobs.resize(8 * (size_t)n_blobs);
    TN_HIP(hipMemcpyAsync(nerf.analytic_blobs.data(), blobs_host, nerf.analytic_blobs.bytes(), hipMemcpyHostToDevice, s));
    OccupancyGrid& og = nerf.occupancy_grid;
    og.set_aabb_scale(aabb);
    og.initialize_if_needed(s);
    TN_TRY(tn_blob_density_grid(og.view(), nerf.analytic_blobs.data(), n_blobs, as_tn(s)));
    TN_TRY(tn_occ_update_bits(og.view(), threshold, as_tn(s)));
    const BoundingBox box(aabb);
    proxy->training_bbox.set(box);
    proxy->render_bbox.set(box);
    proxy->can_render = true;
    TN_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ training
NeRFTrainingController::NeRFTrainingController(NeRFProxy* p, TrainingConfig cfg) : proxy(p), config(cfg) {
    if (!proxy) throw std::runtime_error("NeRFTrainingController: null NeRF");
}

NeRFTrainingController::~NeRFTrainingController() {
    ctx_.reset();
    if (ev_in_) (void)hipEventDestroy(ev_in_);
    if (ev_out_) (void)hipEventDestroy(ev_out_);
}

void NeRFTrainingController::begin_work() {
    if (!ev_in_) TN_HIP(hipEventCreateWithFlags(&ev_in_, hipEventDisableTiming));
    TN_HIP(hipEventRecord(ev_in_, DeviceManager::stream()));
    TN_HIP(hipStreamWaitEvent(ctx_->stream, ev_in_, 0));
}

void NeRFTrainingController::end_work() {
    if (!ev_out_) TN_HIP(hipEventCreateWithFlags(&ev_out_, hipEventDisableTiming));
    TN_HIP(hipEventRecord(ev_out_, ctx_->stream));
    TN_HIP(hipStreamWaitEvent(DeviceManager::stream(), ev_out_, 0));
}

void NeRFTrainingController::setup_data(uint32_t batch_size) {
    if (!proxy->dataset) throw std::runtime_error("setup_data: the NeRF has no dataset attached");
    const hipStream_t s = DeviceManager::stream();
    NeRF& nerf = proxy->nerfs[0];
    const Dataset& ds = *proxy->dataset;
    if (ctx_ && generation_ != proxy->generation) ctx_.reset();  // the slot was destroyed and reused
    if (!ctx_) ctx_ = std::make_unique<Trainer::Context>(&ds, &nerf, batch_size);
    generation_ = proxy->generation;
    ctx_->dataset = &ds;
    ctx_->nerf = &nerf;
    ctx_->batch_size = batch_size;
    ctx_->rng_seed = config.rng_seed;
    ctx_->distortion_lambda = config.distortion_lambda;
    ctx_->occupancy_threshold = config.occupancy_threshold;
    nerf.network.config = config.network;
    const int aabb = std::max(1, (int)ds.bounding_box.size());
    nerf.network.update_aabb_scale_if_needed(aabb);
    nerf.occupancy_grid.set_aabb_scale(ds.bounding_box.size());
    nerf.occupancy_grid.initialize_if_needed(s);
    const uint32_t R = config.ray_capacity ? config.ray_capacity : batch_size;
    TN_HIP(hipStreamSynchronize(ctx_->gen_stream));
    ctx_->enlarge(batch_size, R);
    ctx_->cur = ctx_->last = 0;
    ctx_->prefetched = false;
    const uint32_t init = std::min(ctx_->initial_rays, R);
    TN_HIP(hipMemcpyAsync(ctx_->n_rays_next.data(), &init, sizeof(init), hipMemcpyHostToDevice, s));
    ctx_->samples_total.resize(1);
    ctx_->samples_total.zero(s);
    ctx_->d_rng.resize(1);
    ctx_->d_rng.zero(s);
    ctx_->d_step.resize(1);
    ctx_->d_step.zero(s);
    ctx_->occ_rng_offset = 0;
    ctx_->destroy_graphs();
    // one appearance embedding per training image when the network has them
    // (reference nerf-training-controller.cu:55, nerf-proxy.cuh:106)
    nerf.network.update_appearance_embedding_if_needed((uint32_t)ds.n_images());
    nerf.network.prepare_workspace(batch_size, R);
    nerf.network.update_params_if_needed(s, nerf.params);
    proxy->training_step = 0;
    TN_HIP(hipStreamSynchronize(s));
}

void NeRFTrainingController::teardown() {
    ctx_.reset();
    alpha_selection_probability = 1.0f;
    alpha_selection_threshold = 1.0f;
    min_step_size = NeRFConstants::min_step_size;
    use_distortion_loss = false;
}

void NeRFTrainingController::reset_training() {
    NeRF& nerf = proxy->nerfs[0];
    nerf.params.free_allocations();
    nerf.occupancy_grid.free_device_memory();
    const uint32_t B = ctx_ ? ctx_->batch_size : NeRFConstants::batch_size;
    ctx_.reset();
    setup_data(B);
}

void NeRFTrainingController::load_images(std::function<void(int, int)> on_image_loaded) {
    if (!proxy->dataset) throw std::runtime_error("load_images: no dataset");
    Dataset& ds = *proxy->dataset;
    // a file-backed dataset decodes its images now and drops the CPU copies
    // once they are on the device (reference nerf-training-controller.cu:122-183)
    const bool from_files = !ds.is_loaded_cpu() && !ds.image_paths.empty();
    if (from_files) ds.load_images();
    if (!ds.is_loaded_cpu()) throw std::runtime_error("load_images: the dataset holds no images");
    NeRF& nerf = proxy->nerfs[0];
    nerf.dataset_ws.upload(ds, DeviceManager::stream());
    if (from_files) ds.unload_images();
    nerf.is_image_data_loaded = true;
    proxy->is_dataset_dirty = false;
    if (on_image_loaded) {
        const int n = (int)proxy->dataset->n_images();
        for (int i = 1; i <= n; ++i) on_image_loaded(i, n);
    }
}

void NeRFTrainingController::sync_settings() {
    if (!ctx_) throw std::runtime_error("train_step: call setup_data() first");
    if (generation_ != proxy->generation || !proxy->is_valid)
        throw std::runtime_error("train_step: the NeRF was destroyed since setup_data(); call setup_data() again");
    if (!proxy->nerfs[0].is_image_data_loaded) throw std::runtime_error("train_step: call load_images() first");
    ctx_->alpha_selection_probability = alpha_selection_probability;
    ctx_->alpha_selection_threshold = alpha_selection_threshold;
    ctx_->min_step_size = min_step_size;
    ctx_->use_distortion_loss = use_distortion_loss;
    ctx_->distortion_lambda = config.distortion_lambda;
}

NeRFTrainingController::TrainingMetrics NeRFTrainingController::train_step() {
    sync_settings();
    begin_work();
    const float loss = trainer_.train_step(*ctx_, proxy->training_step);
    end_work();
    proxy->training_step++;
    TrainingMetrics m;
    m.loss = loss;
    m.step = proxy->training_step;
    m.n_rays = ctx_->n_rays_in_batch;
    m.n_samples = ctx_->n_samples_in_batch;
    if (!proxy->can_render && proxy->training_step >= 1) proxy->can_render = true;
    return m;
}

NeRFTrainingController::TrainingMetrics NeRFTrainingController::train_steps(uint32_t n, bool read, bool graph) {
    sync_settings();
    begin_work();
    if (graph) {
        trainer_.enqueue_train_steps_graph(*ctx_, proxy->training_step, n);
        proxy->training_step += n;
    } else {
        for (uint32_t k = 0; k < n; ++k) {
            trainer_.enqueue_train_step(*ctx_, proxy->training_step);
            proxy->training_step++;
        }
    }
    end_work();
    TrainingMetrics m;
    m.step = proxy->training_step;
    if (read) {
        trainer_.read_metrics(*ctx_, m.loss);
        m.n_rays = ctx_->n_rays_in_batch;
        m.n_samples = ctx_->n_samples_in_batch;
    }
    if (proxy->training_step >= 1) proxy->can_render = true;
    return m;
}

NeRFTrainingController::OccupancyGridMetrics NeRFTrainingController::update_occupancy_grid(uint32_t training_step) {
    sync_settings();
    OccupancyGridMetrics m;
    begin_work();
    m.n_occupied = trainer_.update_occupancy_grid(*ctx_, training_step);
    end_work();
    m.n_total = proxy->nerfs[0].occupancy_grid.get_n_total_elements();
    return m;
}

void NeRFTrainingController::set_occupancy_from_density(const float* density_host, float threshold) {
    NeRF& nerf = proxy->nerfs[0];
    OccupancyGrid& og = nerf.occupancy_grid;
    const hipStream_t s = DeviceManager::stream();
    og.initialize_if_needed(s);
    TN_HIP(hipMemcpyAsync(og.density.data(), density_host, og.density.bytes(), hipMemcpyHostToDevice, s));
    TN_TRY(tn_occ_update_bits(og.view(), threshold, as_tn(s)));
    TN_HIP(hipStreamSynchronize(s));
}

std::vector<float> NeRFTrainingController::get_params() const {
    const NeRF& nerf = proxy->nerfs[0];
    std::vector<float> h(nerf.params.params_fp.size());
    // training work ends with the device stream waiting for it
    TN_HIP(hipStreamSynchronize(DeviceManager::stream()));
    TN_HIP(hipMemcpy(h.data(), nerf.params.params_fp.data(), h.size() * sizeof(float), hipMemcpyDeviceToHost));
    return h;
}

void NeRFTrainingController::set_params(const std::vector<float>& h) {
    NeRF& nerf = proxy->nerfs[0];
    if (h.size() != nerf.params.params_fp.size()) throw std::runtime_error("set_params: size mismatch");
    const hipStream_t s = DeviceManager::stream();
    TN_HIP(hipMemcpyAsync(nerf.params.params_fp.data(), h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s));
    if (nerf.params.params_lp.size())
        TN_TRY(tn_cast_params(h.size(), nerf.params.params_fp.data(), 1, nerf.params.params_lp.data(), as_tn(s)));
    TN_HIP(hipStreamSynchronize(s));
}

void NeRFTrainingController::set_timing(bool on) {
    if (!ctx_) throw std::runtime_error("set_timing: call setup_data() first");
    ctx_->timer.enabled = on;
}

std::map<std::string, std::pair<double, uint64_t>> NeRFTrainingController::phase_timings() {
    if (!ctx_) return {};
    ctx_->timer.resolve();
    return ctx_->timer.totals;
}

void NeRFTrainingController::reset_timings() {
    if (ctx_) ctx_->timer.reset();
}

uint64_t NeRFTrainingController::samples_trained() {
    if (!ctx_) return 0;
    unsigned long long h = 0;
    const hipStream_t s = ctx_->stream;
    TN_HIP(hipMemcpyAsync(&h, ctx_->samples_total.data(), sizeof(h), hipMemcpyDeviceToHost, s));
    TN_HIP(hipStreamSynchronize(s));
    return h;
}

void NeRFTrainingController::synchronize() {
    if (ctx_) {
        TN_HIP(hipStreamSynchronize(ctx_->gen_stream));
        TN_HIP(hipStreamSynchronize(ctx_->stream));
    }
    TN_HIP(hipStreamSynchronize(DeviceManager::stream()));
}

// ------------------------------------------------------------------ several NeRFs
NeRFMultiTrainingController::NeRFMultiTrainingController(std::vector<NeRFTrainingController*> trainers,
                                                         bool one_launch_)
    : one_launch(one_launch_), trainers_(std::move(trainers)) {
    if (trainers_.empty()) throw std::runtime_error("NeRFMultiTrainingController: no trainers");
    for (size_t i = 0; i < trainers_.size(); ++i) {
        if (!trainers_[i]) throw std::runtime_error("NeRFMultiTrainingController: null trainer");
        for (size_t j = 0; j < i; ++j)
            if (trainers_[j] == trainers_[i] || trainers_[j]->proxy == trainers_[i]->proxy)
                throw std::runtime_error("NeRFMultiTrainingController: a NeRF appears twice");
    }
    TN_HIP(hipStreamCreateWithFlags(&root_, hipStreamNonBlocking));
    ev_pre_.resize(3 * trainers_.size());
    for (auto& e : ev_pre_) TN_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

NeRFMultiTrainingController::~NeRFMultiTrainingController() {
    (void)hipStreamSynchronize(root_);
    for (auto& g : graph_)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto e : ev_pre_) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(root_);
}

uint64_t NeRFMultiTrainingController::samples_trained() {
    uint64_t s = 0;
    for (auto* t : trainers_) s += t->samples_trained();
    return s;
}

void NeRFMultiTrainingController::synchronize() {
    TN_HIP(hipStreamSynchronize(root_));
    for (auto* t : trainers_) t->synchronize();
}

// The round's graph: every NeRF's training of its current batch and
// generation of its next batch, each captured on its own stream and added to
// one parent graph as an independent child-graph node.
void NeRFMultiTrainingController::capture(int slot, const std::vector<uint64_t>& key) {
    if (graph_[slot]) (void)hipGraphExecDestroy(graph_[slot]);
    graph_[slot] = nullptr;
    hipGraph_t parent = nullptr;
    TN_HIP(hipGraphCreate(&parent, 0));
    for (auto* t : trainers_) t->trainer_.prepare_plans(*t->ctx_);
    try {
        for (auto* t : trainers_) {
            Trainer::Context& ctx = *t->ctx_;
            const int c = ctx.cur;
            for (hipGraph_t child : {t->trainer_.capture_train(ctx, c), t->trainer_.capture_gen(ctx, c ^ 1)}) {
                hipGraphNode_t node = nullptr;
                const hipError_t e = hipGraphAddChildGraphNode(&node, parent, nullptr, 0, child);  // copies `child`
                (void)hipGraphDestroy(child);
                TN_HIP(e);
            }
        }
    } catch (...) {
        (void)hipGraphDestroy(parent);
        throw;
    }
    const hipError_t e = hipGraphInstantiate(&graph_[slot], parent, nullptr, nullptr, 0);
    (void)hipGraphDestroy(parent);
    TN_HIP(e);
    key_[slot] = key;
    ++n_captures;
}

void NeRFMultiTrainingController::train_steps(uint32_t n_rounds) {
    if (n_rounds == 0) return;
    if (one_launch) {
        train_steps_one_launch(n_rounds);
        return;
    }
    // every NeRF replays its own captured step graphs on its own streams,
    // the NeRFs in turn within a round
    std::vector<uint32_t> caps0;
    for (auto* t : trainers_) {
        t->sync_settings();
        t->begin_work();
        t->trainer_.begin_graph_steps(*t->ctx_, t->proxy->training_step);
        caps0.push_back(t->ctx_->n_captures);
    }
    for (uint32_t r = 0; r < n_rounds; ++r)
        for (auto* t : trainers_) {
            t->trainer_.enqueue_graph_step(*t->ctx_);
            t->proxy->training_step++;
            if (!t->proxy->can_render) t->proxy->can_render = true;
        }
    for (size_t k = 0; k < trainers_.size(); ++k) {
        trainers_[k]->end_work();
        n_captures += trainers_[k]->ctx_->n_captures - caps0[k];
    }
}

void NeRFMultiTrainingController::train_steps_one_launch(uint32_t n_rounds) {
    for (size_t k = 0; k < trainers_.size(); ++k) {
        NeRFTrainingController& t = *trainers_[k];
        t.sync_settings();
        t.begin_work();
        Trainer::Context& ctx = *t.ctx_;
        TN_HIP(hipStreamSynchronize(ctx.stream));  // the pinned step value below is free
        TN_HIP(hipStreamSynchronize(root_));
        ctx.nerf->network.prepare_workspace(ctx.batches[0].batch_size, ctx.batches[0].ray_capacity);
        t.trainer_.prepare_plans(ctx);
        const int c = ctx.cur;
        if (!ctx.prefetched) {
            TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ctx.ev_grid, 0));
            TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ctx.ev_consumed[c], 0));
            t.trainer_.generate_next_training_batch(ctx, ctx.batches[c], ctx.gen_stream, &ctx.timer);
            TN_HIP(hipEventRecord(ctx.ev_ready[c], ctx.gen_stream));
            ctx.prefetched = true;
        }
        // the graph follows everything already on this NeRF's streams
        TN_HIP(hipEventRecord(ev_pre_[3 * k], ctx.stream));
        TN_HIP(hipEventRecord(ev_pre_[3 * k + 1], ctx.gen_stream));
        TN_HIP(hipStreamWaitEvent(root_, ev_pre_[3 * k], 0));
        TN_HIP(hipStreamWaitEvent(root_, ev_pre_[3 * k + 1], 0));
        TN_HIP(hipStreamWaitEvent(root_, ctx.ev_ready[c], 0));
        ctx.step_host[0] = t.proxy->training_step;
        TN_HIP(hipMemcpyAsync(ctx.d_step.data(), ctx.step_host.data(), sizeof(uint32_t), hipMemcpyHostToDevice,
                              root_));
    }
    for (uint32_t r = 0; r < n_rounds; ++r) {
        // what every NeRF's captured step has baked in, with the batch each trains on
        std::vector<uint64_t> key;
        for (auto* t : trainers_) {
            const std::vector<uint64_t> kt = t->trainer_.graph_key(*t->ctx_);
            key.insert(key.end(), kt.begin(), kt.end());
            key.push_back((uint64_t)t->ctx_->cur);
        }
        const int slot = trainers_[0]->ctx_->cur;
        if (!graph_[slot] || key_[slot] != key) capture(slot, key);
        TN_HIP(hipGraphLaunch(graph_[slot], root_));
        for (auto* t : trainers_) {
            Trainer::Context& ctx = *t->ctx_;
            const int c = ctx.cur, nx = c ^ 1;
            TN_HIP(hipEventRecord(ctx.ev_ready[nx], root_));
            TN_HIP(hipEventRecord(ctx.ev_consumed[c], root_));
            ctx.last = c;
            ctx.cur = nx;
            ctx.prefetched = true;
            t->proxy->training_step++;
            if (!t->proxy->can_render) t->proxy->can_render = true;
        }
    }
    // later work on each NeRF's own streams follows the rounds
    for (size_t k = 0; k < trainers_.size(); ++k) {
        Trainer::Context& ctx = *trainers_[k]->ctx_;
        TN_HIP(hipEventRecord(ev_pre_[3 * k + 2], root_));
        TN_HIP(hipStreamWaitEvent(ctx.stream, ev_pre_[3 * k + 2], 0));
        TN_HIP(hipStreamWaitEvent(ctx.gen_stream, ev_pre_[3 * k + 2], 0));
        trainers_[k]->end_work();
    }
}

std::vector<size_t> NeRFTrainingController::get_cuda_memory_allocated() const {
    size_t total = 0;
    if (ctx_) total += ctx_->bytes();
    const NeRF& nerf = proxy->nerfs[0];
    total += nerf.network.workspace.bytes();
    total += nerf.params.params_fp.bytes() + nerf.params.params_lp.bytes() + nerf.params.gradients.bytes() +
             nerf.params.m1.bytes() + nerf.params.m2.bytes() + nerf.params.steps.bytes();
    total += nerf.occupancy_grid.bitfield.bytes() + nerf.occupancy_grid.density.bytes();
    return {total};
}

}  // namespace turbo
