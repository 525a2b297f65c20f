// Public C++ API for training (and the NeRF slot manager), mirroring the
// reference's NeRFTrainingController (src/controllers/nerf-training-controller.h:23-89),
// NeRFProxy (src/models/nerf-proxy.cuh:20-141) and NeRFManager
// (src/services/nerf-manager.cuh:27-183).
#pragma once

#include <any>
#include <functional>
#include <map>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "core.hpp"

namespace turbo {

struct NeRFProxy {
    std::vector<NeRF> nerfs;  // one per device (one device here)
    std::optional<Dataset> dataset;
    UpdatableProperty<BoundingBox> render_bbox;
    UpdatableProperty<BoundingBox> training_bbox;
    UpdatableProperty<Transform4f> transform;
    bool is_visible = true;
    bool is_dataset_dirty = true;
    bool can_render = false;
    uint32_t training_step = 0;
    int id = -1;
    bool is_valid = false;
    NeRFProxy* clone_source = nullptr;
    // Bumped whenever the slot is destroyed or handed out again; a controller
    // remembers the value it was set up with and refuses to run on a newer one.
    uint64_t generation = 0;

    NeRFProxy() : nerfs(1) {}
    NeRFProxy(const NeRFProxy&) = delete;
    NeRFProxy& operator=(const NeRFProxy&) = delete;
    // Return the slot to a fresh state in place: the NeRF objects are reset
    // element by element, so pointers to nerfs[i] held elsewhere keep
    // pointing at live (now empty) objects.
    void reset() {
        for (auto& n : nerfs) n = NeRF();
        dataset.reset();
        render_bbox = {};
        training_bbox = {};
        transform = {};
        is_visible = true;
        is_dataset_dirty = true;
        can_render = false;
        training_step = 0;
        clone_source = nullptr;
        ++generation;
    }
    void attach_dataset(const Dataset& ds) {
        dataset = ds;
        training_bbox.set(ds.bounding_box);
        render_bbox.set(ds.bounding_box);
        is_dataset_dirty = true;
    }
    void detach_dataset() { dataset.reset(); }
    bool can_train() const { return dataset.has_value(); }
    bool is_image_data_loaded() const { return !nerfs.empty() && nerfs[0].is_image_data_loaded; }
    bool is_dirty() const { return is_dataset_dirty || training_bbox.is_dirty() || transform.is_dirty(); }
    void free_device_memory() { for (auto& n : nerfs) n.free_device_memory(); }
};

// Make `proxy` a synthetic scene: the renderer evaluates the blob field
// (n_blobs rows of {cx, cy, cz, s, p, r, g, b}, world units) instead of a
// network; the box is a cube of side `aabb`, the occupancy grid (log2(aabb)+1
// cascades of 128^3) is the field's density at cell centres thresholded at
// `threshold`.
void set_analytic_field(NeRFProxy* proxy, const float* blobs_host, uint32_t n_blobs, float aabb, float threshold);

// Fixed slot pool handing out stable raw pointers (reference nerf-manager.cuh:49-51)
class NeRFManager {
public:
    NeRFManager() : proxies_(NeRFConstants::n_max_nerfs) {}
    NeRFProxy* create();
    NeRFProxy* clone(NeRFProxy* src);
    void destroy(NeRFProxy* p);
    std::vector<NeRFProxy*> get_proxies();

private:
    std::vector<std::unique_ptr<NeRFProxy>> proxies_;
};

// Runtime configuration beyond the reference's public fields (the reference
// fixes these at compile time; the benchmark configurations change them).
struct TrainingConfig {
    NetworkConfig network;
    uint64_t rng_seed = 0x5eed;
    float distortion_lambda = NeRFConstants::mipNeRF360_distortion_loss_lambda;
    float occupancy_threshold = NeRFConstants::occupancy_threshold;
    uint32_t ray_capacity = 0;  // 0 = batch_size, as the reference sizes ray buffers
};

struct NeRFTrainingController {
    struct TrainingMetrics {
        uint32_t step = 0, n_rays = 0, n_samples = 0;
        float loss = 0.0f;
        std::map<std::string, std::any> as_map() const {
            return {{"step", step}, {"n_rays", n_rays}, {"n_samples", n_samples}, {"loss", loss}};
        }
    };
    struct OccupancyGridMetrics {
        uint32_t n_occupied = 0, n_total = 0;
        std::map<std::string, std::any> as_map() const { return {{"n_occupied", n_occupied}, {"n_total", n_total}}; }
    };

    explicit NeRFTrainingController(NeRFProxy* proxy, TrainingConfig cfg = TrainingConfig());
    ~NeRFTrainingController();

    void setup_data(uint32_t batch_size = NeRFConstants::batch_size);
    void teardown();
    void reset_training();
    void load_images(std::function<void(int, int)> on_image_loaded = {});
    TrainingMetrics train_step();
    // Enqueue n steps without host synchronisation; metrics of the last one.
    // graph: each step is the replay of a captured HIP graph (captured on
    // first use and again whenever a setting or buffer it was built from
    // changes); results are the same as without.
    TrainingMetrics train_steps(uint32_t n_steps, bool read_metrics = true, bool graph = false);
    OccupancyGridMetrics update_occupancy_grid(uint32_t training_step);
    std::vector<size_t> get_cuda_memory_allocated() const;

    // occupancy bitfield from an external density grid (synthetic scenes)
    void set_occupancy_from_density(const float* density_host, float threshold);
    // raw parameter access (fp32 master weights), for tests and snapshots
    std::vector<float> get_params() const;
    void set_params(const std::vector<float>& p);

    // per-phase HIP-event timing of enqueued steps (name -> ms, launches)
    void set_timing(bool on);
    std::map<std::string, std::pair<double, uint64_t>> phase_timings();
    void reset_timings();
    // samples trained on since setup_data (device running total; synchronises)
    uint64_t samples_trained();
    void synchronize();
    // Test support: the training context (batches, streams), or null before
    // setup_data(). The batch the last step trained on is
    // context()->batches[context()->last].
    Trainer::Context* context() { return ctx_.get(); }

    float alpha_selection_threshold = 1.0f;
    float alpha_selection_probability = 1.0f;
    float min_step_size = NeRFConstants::min_step_size;
    bool use_distortion_loss = false;
    NeRFProxy* proxy = nullptr;
    TrainingConfig config;

private:
    friend class NeRFMultiTrainingController;
    std::unique_ptr<Trainer::Context> ctx_;
    uint64_t generation_ = 0;  // proxy->generation when setup_data() ran
    Trainer trainer_;
    hipEvent_t ev_in_ = nullptr, ev_out_ = nullptr;
    void sync_settings();
    // Training runs on the trainer's own streams; begin_work orders it after
    // the work already on the device stream (uploads, parameter writes),
    // end_work orders later device-stream work (renders, snapshots) after it.
    void begin_work();
    void end_work();
};

// Several NeRFs trained side by side (reference: one training controller per
// NeRF, stepped in turn). Each round enqueues one training step of every NeRF
// as replays of captured HIP graphs, so the NeRFs' steps may run
// concurrently; results are those of stepping the controllers one after
// another. By default every NeRF replays its own step graphs on its own
// streams (two launches per NeRF and round); with one_launch the round is
// one launch of a graph holding every NeRF's captured step as an independent
// child graph (slower on this runtime: its child-graph launches leave longer
// gaps between kernels). A controller may still be stepped on its own
// between rounds.
class NeRFMultiTrainingController {
public:
    explicit NeRFMultiTrainingController(std::vector<NeRFTrainingController*> trainers, bool one_launch = false);
    ~NeRFMultiTrainingController();
    NeRFMultiTrainingController(const NeRFMultiTrainingController&) = delete;
    NeRFMultiTrainingController& operator=(const NeRFMultiTrainingController&) = delete;
    void train_steps(uint32_t n_rounds);
    void synchronize();
    uint64_t samples_trained();  // summed over the NeRFs (synchronises)
    uint32_t n_captures = 0;  // graphs captured for the rounds so far
    const bool one_launch;

private:
    std::vector<NeRFTrainingController*> trainers_;
    void train_steps_one_launch(uint32_t n_rounds);
    hipStream_t root_ = nullptr;
    std::vector<hipEvent_t> ev_pre_;
    hipGraphExec_t graph_[2] = {nullptr, nullptr};
    std::vector<uint64_t> key_[2];
    void capture(int slot, const std::vector<uint64_t>& key);
};

}  // namespace turbo
