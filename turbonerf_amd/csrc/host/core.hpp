// Core engines of the training/rendering hot path: OccupancyGrid,
// NerfNetwork and Trainer, with the reference's class and method names
// (src/core/occupancy-grid.cuh, nerf-network.cuh, trainer.cuh). Every device
// computation is a call into the tn_* kernel library.
#pragma once

#include <cmath>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "models.hpp"

namespace turbo {

// reference src/utils/nerf-constants.cuh
namespace NeRFConstants {
constexpr float min_step_size = 0.00169145586f;  // sqrt(3)/1024
constexpr float cone_angle = 1.0f / 256.0f;
constexpr uint32_t batch_size = 2 << 21;
constexpr float min_transmittance = 1e-4f;
constexpr float occupancy_decay = 0.95f;
constexpr float mipNeRF360_distortion_loss_lambda = 0.01f;
constexpr float occupancy_threshold = 0.01f * min_step_size;
constexpr uint32_t n_steps_per_render_compaction = 64;
constexpr float learning_rate_decay = 3.3e-5f;
constexpr uint32_t n_max_nerfs = 2 << 10;
}  // namespace NeRFConstants

// Optional per-phase timing with HIP events on the work stream. Records are
// resolved (and the events recycled) by resolve(), which synchronises.
class PhaseTimer {
public:
    bool enabled = false;
    ~PhaseTimer();
    void begin(const char* name, hipStream_t s);
    void end(hipStream_t s);
    void resolve();
    void reset() { resolve(); totals.clear(); }
    // name -> (total milliseconds, launches)
    std::map<std::string, std::pair<double, uint64_t>> totals;

private:
    struct Rec { std::string name; hipEvent_t a, b; };
    std::vector<Rec> pending_;
    std::vector<hipEvent_t> pool_;
    hipEvent_t get_();
};

struct ScopedPhase {
    PhaseTimer* t;
    hipStream_t s;
    ScopedPhase(PhaseTimer* t_, const char* name, hipStream_t s_) : t(t_), s(s_) {
        if (t && t->enabled) t->begin(name, s);
    }
    ~ScopedPhase() {
        if (t && t->enabled) t->end(s);
    }
};

// Multi-cascade occupancy grid: Morton-ordered bitfield (bit k = cascade k)
// and per-cascade fp32 densities (reference src/core/occupancy-grid.cuh:22-287).
struct OccupancyGrid {
    int n_levels = 1;
    int resolution = 128;
    DeviceBuffer<uint8_t> bitfield;
    DeviceBuffer<float> density;
    DeviceBuffer<uint8_t> workspace;  // popcount partials
    DeviceBuffer<uint32_t> count;

    static int get_max_n_levels(float aabb_size) { return (int)std::log2(aabb_size) + 1; }
    uint32_t volume() const { return (uint32_t)resolution * resolution * resolution; }
    void set_aabb_scale(float aabb) { n_levels = get_max_n_levels(aabb); }

    // allocate and set every bit (all cells occupied), densities zero
    void initialize_if_needed(hipStream_t s) {
        const size_t vol = volume();
        if (bitfield.size() == vol && density.size() == vol * (size_t)n_levels) return;
        bitfield.resize(vol);
        density.resize(vol * (size_t)n_levels);
        bitfield.fill_bytes(0xFF, s);
        density.zero(s);
    }
    void set_bitfield(int value, hipStream_t s) { bitfield.fill_bytes(value, s); }
    tn_occupancy_grid_t view() const { return tn_occupancy_grid_t{bitfield.data(), density.data(), n_levels, resolution}; }
    uint32_t get_n_total_elements() const { return (uint32_t)n_levels * volume(); }
    // number of set bits (device popcount), synchronises the stream
    uint32_t count_occupied(hipStream_t s);
    void free_device_memory() { bitfield.release(); density.release(); workspace.release(); count.release(); }
};

// Network and optimiser configuration. Defaults follow the reference
// (src/core/nerf-network.cu:86-121,175-189); beta2 defaults to the tcnn
// value the reference leaves in place (0.999).
struct NetworkConfig {
    int precision = 1;  // 0 = plain fp32, 1 = fp16 MFMA with fp32 accumulation and master weights
    int n_levels = 16;
    int n_features = 2;
    int log2_hashmap_size = 19;
    int base_resolution = 16;
    float finest_resolution_per_unit = 2048.0f;  // N_max = 2048 * aabb
    float hash_init_scale = 1e-4f;
    uint64_t seed = 72791;
    float learning_rate = 1e-2f;
    float learning_rate_decay = NeRFConstants::learning_rate_decay;
    float beta1 = 0.9f;
    float beta2 = 0.999f;
    float epsilon = 1e-15f;
    float l2_reg = 1e-6f;
    // appearance embedding width: 0 (none; the benchmark network) or
    // TN_APPEARANCE_DIMS per training image, as the reference always trains
    // (src/core/nerf-network.cu:27-72, nerf-constants.cuh:31)
    int appearance_dims = 0;
    // mixed precision: inference (rendering, the occupancy update) encodes to
    // fp16 features and runs the network on inputs rounded to fp16 once, as
    // the reference's fp16 network does; false: the training forward's fp32
    // features and split fp16 first-layer inputs
    bool inference_fp16_features = true;
    float loss_scale() const { return precision == 1 ? 1024.0f : 1.0f; }
};

// Flat parameters (hash table, density MLP, colour MLP, then with an
// appearance embedding the colour MLP's 16 embedding columns W3e [64][16] and
// the embeddings [n_appearances][16]), fp32 master copy, optional fp16 copy,
// fp32 gradients and Adam state (reference
// src/workspaces/network-params-workspace.cuh, src/core/adam-optimizer.cuh).
struct NetworkParamsWorkspace {
    uint32_t n_grid = 0, n_total = 0, n_appearances = 0;
    uint32_t w3e_offset() const { return n_grid + TN_DENSITY_MLP_PARAMS + TN_COLOR_MLP_PARAMS; }
    uint32_t emb_offset() const { return w3e_offset() + TN_APPEARANCE_WEIGHTS; }
    // end of the network weights (L2-regularised by Adam); the embeddings follow
    uint32_t n_net_end() const { return n_appearances ? emb_offset() : n_total; }
    DeviceBuffer<float> params_fp;
    DeviceBuffer<uint8_t> params_lp;  // fp16 copy when training in mixed precision
    DeviceBuffer<float> gradients;
    DeviceBuffer<float> m1, m2;
    DeviceBuffer<uint32_t> steps;
    void free_allocations() {
        params_fp.release(); params_lp.release(); gradients.release(); m1.release(); m2.release(); steps.release();
        n_total = 0;
    }
};

// Per-batch buffers of the network pass (reference NetworkWorkspace,
// src/workspaces/network-workspace.cuh:43-95).
struct NetworkWorkspace {
    uint32_t batch_size = 0, ray_capacity = 0;
    int precision = -1;
    DeviceBuffer<uint8_t> features, dfeat;         // fp32 [32][B]
    DeviceBuffer<uint8_t> density, rgb, dL_drgb;   // network precision
    DeviceBuffer<float> dL_ds, dL_dist, ray_loss, ray_dist_loss, loss_out;
    DeviceBuffer<uint8_t> composite_scratch, distortion_scratch, mlp_ws, hash_ws;
    DeviceBuffer<float> app_dz, app_ws;  // appearance embedding: colour-layer dz [64][B], per-image sums
    DeviceBuffer<uint32_t> info_tail;    // batch descriptor of the second half (tn_batch_tail)
    void enlarge(uint32_t B, uint32_t R, int precision);
    size_t bytes() const;
};

struct NerfNetwork {
    struct Settings {
        bool use_distortion_loss = false;
        float distortion_lambda = NeRFConstants::mipNeRF360_distortion_loss_lambda;
        float min_transmittance = NeRFConstants::min_transmittance;
        PhaseTimer* timer = nullptr;
        unsigned long long* samples_total = nullptr;  // device counter of trained samples, if kept
        // device training step: incremented at the end of the step when
        // given; with step_for_lr also the step the learning-rate schedule
        // reads (a captured step replays with it advancing on the device)
        uint32_t* step_counter = nullptr;
        bool step_for_lr = false;
        // the batch's hash-grid backward plan (tn_hashgrid_backward_plan),
        // formed ahead from the positions; null: the backward forms it
        const void* hash_plan = nullptr;
        size_t hash_plan_bytes = 0;
        // with a plan and TN_TUNE_STEP_HALVES: a second stream and four
        // events for running the batch's two halves side by side
        hipStream_t aux = nullptr;
        const hipEvent_t* split_events = nullptr;
    };

    NetworkConfig config;
    tn_hashgrid_t grid{};
    int aabb_scale = 0;
    NetworkWorkspace workspace;
    // appearance embedding: one per training image (0: none), and the
    // per-image bias rows W3e emb[k] [n_appearances][64] of the current
    // parameters
    uint32_t n_appearances = 0;
    DeviceBuffer<float> app_bias;

    // an embedding for n appearances, or none for 0 (reference
    // update_appearance_embedding_if_needed, src/core/nerf-network.cu:27-72);
    // the parameters are laid out again on the next update_params_if_needed
    void update_appearance_embedding_if_needed(uint32_t n) { n_appearances = config.appearance_dims ? n : 0; }

    // (re)build the encoding layout for a scene of side aabb (reference
    // update_aabb_scale_if_needed, src/core/nerf-network.cu:77-124)
    void update_aabb_scale_if_needed(int aabb);
    // allocate and seed parameters, Adam state (reference :127-190)
    void update_params_if_needed(hipStream_t s, NetworkParamsWorkspace& p);
    // every buffer a training step of this batch shape uses, allocated up
    // front (a captured step may not allocate)
    void prepare_workspace(uint32_t batch_size, uint32_t ray_capacity);
    const void* params_for_kernels(const NetworkParamsWorkspace& p) const {
        return config.precision == 1 ? (const void*)p.params_lp.data() : (const void*)p.params_fp.data();
    }

    // One training step over a batch already generated on the device. The
    // losses (reconstruction, distortion) are left in workspace.loss_out.
    // Reference NerfNetwork::train (src/core/nerf-network.cu:223-317).
    // appearance_id_batch: per-sample image (non-decreasing over the batch;
    // used with an appearance embedding only).
    void train(hipStream_t s, NetworkParamsWorkspace& p, uint32_t step, uint32_t batch_size, uint32_t ray_capacity,
               const uint32_t* info, const float* random_rgb, const uint32_t* ray_steps, const uint32_t* ray_offset,
               const uint32_t* appearance_id_batch, const float* t_batch, const float* pos_batch,
               const float* dir_batch, const float* dt_batch, const float* m_norm_batch, const float* dt_norm_batch,
               const float* target_rgba, const Settings& settings);

    // density (and colour) of n points (reference NerfNetwork::inference,
    // src/core/nerf-network.cu:319-429); outputs in the network precision.
    // With an appearance embedding every sample takes appearance 0, as the
    // reference's renderer sets (src/core/renderer.cu:146-153).
    void inference(hipStream_t s, const NetworkParamsWorkspace& p, uint32_t n, const uint32_t* info, uint32_t stride,
                   const float* pos_batch, const float* dir_batch, void* features, void* density_out, void* rgb_out,
                   bool use_color_network);
    size_t elem_bytes() const { return config.precision == 1 ? 2 : 4; }
};

// Per-step training buffers (reference TrainingWorkspace,
// src/workspaces/training-workspace.cuh:20-113).
struct TrainingWorkspace {
    uint32_t batch_size = 0, ray_capacity = 0;
    DeviceBuffer<float> random_float;
    DeviceBuffer<float> pix_rgba, ray_origin, ray_dir, ray_t, ray_t_max;
    DeviceBuffer<uint32_t> ray_img_id, ray_step, ray_offset;
    DeviceBuffer<uint8_t> ray_alive;
    DeviceBuffer<uint32_t> sample_img_id;
    DeviceBuffer<float> sample_t, sample_pos, sample_dir, sample_dt, sample_m_norm, sample_dt_norm;
    DeviceBuffer<uint32_t> info;
    DeviceBuffer<uint8_t> scan_ws;
    // the hash-grid backward's plan for this batch (record counts and offsets
    // per bin, from the sample positions), formed with the batch
    DeviceBuffer<uint8_t> hash_plan;
    void enlarge(uint32_t B, uint32_t R);
    size_t bytes() const;
    void free_allocations();
};

// Device-resident dataset (reference DatasetWorkspace).
struct DatasetWorkspace {
    DeviceBuffer<uint8_t> cameras;  // tn_camera_t[n_images]
    DeviceBuffer<uint8_t> images;   // RGBA8
    DeviceBuffer<uint8_t> bbox;     // tn_bbox_t
    uint32_t n_images = 0;
    int2v image_dims;
    uint32_t n_pixels_per_image = 0;
    void upload(const Dataset& ds, hipStream_t s);
};

// One NeRF's device state (reference src/models/nerf.cuh:17-32).
struct NeRF {
    NerfNetwork network;
    NetworkParamsWorkspace params;
    OccupancyGrid occupancy_grid;
    DatasetWorkspace dataset_ws;
    bool is_image_data_loaded = false;
    // Synthetic scene: a Gaussian-blob field (rows {cx, cy, cz, s, p, r, g, b}
    // in world units) that the renderer evaluates in place of the network.
    DeviceBuffer<float> analytic_blobs;
    bool holds_device_memory() const {
        return params.params_fp.size() || occupancy_grid.bitfield.size() || dataset_ws.images.size() ||
               network.workspace.bytes();
    }
    void free_device_memory() {
        params.free_allocations();
        occupancy_grid.free_device_memory();
        analytic_blobs.release();
    }
};

struct Trainer {
    // Batch generation (rays, marching, scan, sample generation) needs only
    // the occupancy grid and the dataset, not the network, so it runs one
    // step ahead on its own stream into the other of two batch workspaces
    // while the network trains on the current one. An occupancy update
    // discards the batch generated ahead, so every batch is marched through
    // the grid the reference would use.
    struct Context {
        hipStream_t stream;      // network, optimiser and occupancy update (the trainer's own)
        hipStream_t gen_stream;  // batch generation
        // the second half of a batch's per-sample network passes, forked
        // from and joined back into `stream` within each step
        hipStream_t aux_stream;
        hipEvent_t ev_split[4]{};
        TrainingWorkspace batches[2];
        int cur = 0;              // batch the next step trains on
        int last = 0;             // batch the last step trained on
        bool prefetched = false;  // batches[cur] already holds a generated batch
        hipEvent_t ev_ready[2]{}, ev_consumed[2]{}, ev_grid{}, ev_gen_idle{};
        DeviceBuffer<uint32_t> n_rays_next;  // device-side ray count for the next batch
        // position of the batch-generation uniform stream and the training
        // step, both on the device, so a captured step needs no host values
        DeviceBuffer<unsigned long long> d_rng;
        DeviceBuffer<uint32_t> d_step;
        PinnedBuffer<uint32_t> step_host{1};
        uint64_t occ_rng_offset = 0;  // occupancy-update uniforms (their own key)
        // captured steps: train_graph[c] trains on batches[c] (network
        // stream), gen_graph[c] generates batches[c] (generation stream); the
        // two streams order them with the events of eager steps, so a step's
        // training overlaps the next batch's generation as in eager steps.
        // Recaptured when anything they were built from changes.
        hipGraphExec_t train_graph[2] = {nullptr, nullptr};
        hipGraphExec_t gen_graph[2] = {nullptr, nullptr};
        std::vector<uint64_t> graph_key;
        uint32_t n_captures = 0;  // graphs captured so far
        void destroy_graphs();
        const Dataset* dataset;
        NeRF* nerf;
        uint32_t batch_size;
        uint32_t n_rays_in_batch = 0;
        uint32_t n_samples_in_batch = 0;
        float n_rays_per_image = 0;
        float alpha_selection_threshold = 1.0f;
        float alpha_selection_probability = 1.0f;
        float min_step_size = NeRFConstants::min_step_size;
        bool use_distortion_loss = false;
        float distortion_lambda = NeRFConstants::mipNeRF360_distortion_loss_lambda;
        float occupancy_threshold = NeRFConstants::occupancy_threshold;
        uint64_t rng_seed = 0x5eed;
        uint32_t initial_rays = 16384;
        PinnedBuffer<uint32_t> info_host{4};
        PinnedBuffer<float> loss_host{2};
        PinnedBuffer<uint32_t> count_host{1};
        PhaseTimer timer;
        DeviceBuffer<unsigned long long> samples_total;  // running sum of samples trained on

        Context(const Dataset* ds, NeRF* n, uint32_t B);
        ~Context();
        Context(const Context&) = delete;
        Context& operator=(const Context&) = delete;
        void enlarge(uint32_t B, uint32_t R) {
            batches[0].enlarge(B, R);
            batches[1].enlarge(B, R);
            n_rays_next.resize(1);
        }
        size_t bytes() const { return batches[0].bytes() + batches[1].bytes(); }
    };

    // Enqueue one training step; no host synchronisation.
    void enqueue_train_step(Context& ctx, uint32_t training_step);
    // Enqueue n training steps as replays of captured HIP graphs (one graph
    // launch per step, the next batch generated inside it); training steps
    // first_step .. first_step + n - 1. Same results as enqueue_train_step.
    void enqueue_train_steps_graph(Context& ctx, uint32_t first_step, uint32_t n);
    // enqueue_train_steps_graph in two parts, for callers that interleave the
    // steps of several trainers: begin_graph_steps once (workspace sizes, the
    // captured graphs' key, the device step counter; synchronises the
    // trainer's network stream), then enqueue_graph_step once per step.
    void begin_graph_steps(Context& ctx, uint32_t first_step);
    void enqueue_graph_step(Context& ctx);
    // One training step, synchronising like the reference: throws when the
    // batch has no rays/samples, returns the loss (reference trainer.cu:313-349).
    float train_step(Context& ctx, uint32_t training_step);
    // Batch generation into `ws` on `s` (reference trainer.cu:39-217).
    void generate_next_training_batch(Context& ctx, TrainingWorkspace& ws, hipStream_t s, PhaseTimer* tm);
    // Occupancy update; returns the number of occupied bits (reference trainer.cu:221-311).
    uint32_t update_occupancy_grid(Context& ctx, uint32_t training_step);
    // Read back batch counts and losses of the last enqueued step.
    void read_metrics(Context& ctx, float& loss);

    // Building blocks of steps. train_network trains on batches[c] on the
    // network stream (`graph`: it is being captured, so the learning-rate
    // step is read on the device and no phase is timed). capture_train and
    // capture_gen return the captured training of batches[c] and generation
    // into batches[c] (the caller owns the graphs). graph_key lists what a
    // captured step has baked in.
    void train_network(Context& ctx, int c, uint32_t training_step, bool graph);
    hipGraph_t capture_train(Context& ctx, int c);
    hipGraph_t capture_gen(Context& ctx, int c);
    std::vector<uint64_t> graph_key(const Context& ctx) const;
    // Eager step on batches[c] while batches[c ^ 1] is generated.
    void train_batch(Context& ctx, int c, uint32_t training_step);
    // sizes both batches' hash-grid backward plans (before any capture)
    void prepare_plans(Context& ctx);

private:
    void instantiate_graphs(Context& ctx, int c);
};

}  // namespace turbo
