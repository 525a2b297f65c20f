// Rendering API mirroring the reference's NeRFRenderingController
// (src/controllers/nerf-rendering-controller.h:14-41), RenderRequest
// (src/models/render-request.cuh:20-87), CPURenderBuffer
// (src/render-targets/cpu-render-buffer.cuh:11-69) and Renderer
// (src/core/renderer.cuh:36-64).
#pragma once

#include <atomic>
#include <functional>
#include <memory>
#include <vector>

#include "controllers.hpp"

namespace pybind11 {
class module_;
}

namespace turbo {

enum class RenderPattern { LinearChunks = 0, HexagonalGrid = 1, RectangularGrid = 2 };
enum class RenderFlags : int { Preview = 1 << 0, Final = 1 << 1 };
// flag sets (reference src/utils/render-flags.cuh:12-29)
inline RenderFlags operator|(RenderFlags a, RenderFlags b) {
    return static_cast<RenderFlags>(static_cast<int>(a) | static_cast<int>(b));
}
inline RenderFlags operator&(RenderFlags a, RenderFlags b) {
    return static_cast<RenderFlags>(static_cast<int>(a) & static_cast<int>(b));
}
inline RenderFlags& operator|=(RenderFlags& a, RenderFlags b) { return a = a | b; }
inline RenderFlags& operator&=(RenderFlags& a, RenderFlags b) { return a = a & b; }

struct RenderProperties {
    bool show_near_planes = false;
    bool show_far_planes = false;
};
struct RenderModifiers {
    RenderProperties properties;
};

// Output surface: interleaved float RGBA (premultiplied), host memory.
struct RenderTarget {
    int width = 0, height = 0;
    virtual ~RenderTarget() = default;
    virtual void set_size(int w, int h) = 0;
    virtual void free() = 0;
    // The image as an 8-bit RGBA PNG: each channel v becomes (uint8)(v * 255),
    // truncated, values outside [0, 1] clamped (reference render-target.cuh:51-59,
    // gpu-image.cu:13-24,47-83)
    virtual void save_image(const std::string& file_path) = 0;
    int n_pixels() const { return width * height; }
};

struct CPURenderBuffer : RenderTarget {
    std::vector<float> rgba;
    void set_size(int w, int h) override {
        width = w;
        height = h;
        rgba.assign((size_t)w * h * 4, 0.0f);
    }
    void free() override { rgba.clear(); rgba.shrink_to_fit(); width = height = 0; }
    float* get_rgba() { return rgba.data(); }
    void save_image(const std::string& file_path) override;
};

// Output surface in device memory (reference CUDARenderBuffer,
// src/render-targets/cuda-render-buffer.cuh:11-42): renders write straight
// into it, with no copy to the host.
struct CUDARenderBuffer : RenderTarget {
    DeviceBuffer<float> rgba;
    void set_size(int w, int h) override {
        width = w;
        height = h;
        rgba.resize((size_t)w * h * 4);
        rgba.zero(DeviceManager::stream());
        TN_HIP(hipStreamSynchronize(DeviceManager::stream()));
    }
    void free() override { rgba.release(); width = height = 0; }
    float* data() { return rgba.data(); }
    void save_image(const std::string& file_path) override;
};

// Spatial effects applied to a NeRF's samples when it renders (reference
// src/effects/spatial-effect.cuh, spatial/repeater.cuh:14-40).
struct SpatialEffect {
    virtual ~SpatialEffect() = default;
    // the box the effect lets the NeRF occupy, given the box it had
    virtual BoundingBox get_bbox(const BoundingBox& bbox) const = 0;
};

// Repeats the contents of source_bbox to fill extend_bbox: a sample at p is
// evaluated at ((p - min) mod size + size) mod size + min of the source box
// (reference repeater.cu:8-76; its transform is stored but not applied).
struct RepeaterEffect : SpatialEffect {
    BoundingBox source_bbox, extend_bbox;
    Transform4f transform;
    RepeaterEffect(const BoundingBox& src, const BoundingBox& ext, const Transform4f& t)
        : source_bbox(src), extend_bbox(ext), transform(t) {}
    BoundingBox get_bbox(const BoundingBox& b) const override;
};

struct NeRFRenderable {
    NeRFProxy* proxy = nullptr;
    BoundingBox bounding_box;
    std::vector<std::shared_ptr<SpatialEffect>> spatial_effects;
    explicit NeRFRenderable(NeRFProxy* p, std::vector<std::shared_ptr<SpatialEffect>> fx = {})
        : proxy(p), bounding_box(p ? p->render_bbox.get() : BoundingBox()), spatial_effects(std::move(fx)) {}
    NeRFRenderable(NeRFProxy* p, const BoundingBox& b, std::vector<std::shared_ptr<SpatialEffect>> fx = {})
        : proxy(p), bounding_box(b), spatial_effects(std::move(fx)) {}
};

using OnCompleteCallback = std::function<void()>;
using OnProgressCallback = std::function<void(float)>;
using OnCancelCallback = std::function<void()>;

struct RenderRequest {
    Camera camera;
    std::vector<NeRFRenderable> renderables;
    RenderTarget* output = nullptr;
    RenderModifiers modifiers;
    RenderFlags flags = RenderFlags::Final;
    OnCompleteCallback on_complete;
    OnProgressCallback on_progress;
    OnCancelCallback on_cancel;
    std::atomic<bool> canceled{false};
    RenderRequest(const Camera& c, std::vector<NeRFRenderable> r, RenderTarget* out,
                  RenderModifiers m = RenderModifiers(), RenderFlags f = RenderFlags::Final,
                  OnCompleteCallback oc = nullptr, OnProgressCallback op = nullptr, OnCancelCallback ocn = nullptr)
        : camera(c), renderables(std::move(r)), output(out), modifiers(m), flags(f), on_complete(std::move(oc)),
          on_progress(std::move(op)), on_cancel(std::move(ocn)) {}
    void cancel() { canceled = true; }
    bool is_canceled() const { return canceled.load(); }
};

// Device buffers of one render pass.
struct RenderingWorkspace {
    uint32_t ray_capacity = 0, sample_capacity = 0;
    DeviceBuffer<float> ray_ori, ray_dir, ray_t, ray_tmax, rgba, image;
    DeviceBuffer<uint32_t> pixel, n_steps, offset, info;
    DeviceBuffer<uint8_t> alive, scan_ws, to_local;
    DeviceBuffer<uint32_t> s_img;
    DeviceBuffer<float> s_t, s_pos, s_dir, s_dt, s_m, s_dtn;
    DeviceBuffer<uint8_t> features, density, color;
    // several NeRFs (fixed-step world samples shared by all NeRFs)
    DeviceBuffer<float> m_trans, m_pos, m_dt, m_net_pos, m_net_dir, m_net_dt, m_c_pos, m_c_dir, m_rgba;
    DeviceBuffer<uint32_t> m_valid, m_offs, m_count, m_alive_n;
    DeviceBuffer<uint8_t> m_scan_ws, m_nerfs;
    void enlarge(uint32_t rays, uint32_t samples, size_t elem_bytes);
    void enlarge_multi(uint32_t rays, uint32_t samples, uint32_t n_nerfs);
};

struct Renderer {
    struct Settings {
        float min_step_size = NeRFConstants::min_step_size;
        uint32_t max_samples_per_ray = 0xFFFFFFFFu;
        float min_transmittance = NeRFConstants::min_transmittance;
    };
    RenderingWorkspace ws;
    // samples per ray of the last pass, to size the next pass's rays to the
    // sample budget (0: not known yet)
    double samples_per_ray = 0.0;
    // Render pixels [start, start + n) of the request's camera into `image`
    // (device, interleaved RGBA). Returns the number of sample evaluations.
    uint64_t render_pixels(hipStream_t s, const RenderRequest& req, const NeRFRenderable& r, uint32_t start, uint32_t n,
                           float* image, const Settings& st);
    // Several NeRFs (or one with spatial effects) the reference's way
    // (renderer.cu:344-521): every ray steps through world space with the
    // fixed step rule, each NeRF evaluates the samples inside its box on
    // occupied cells, a sample's alphas add (capped at 1) and its colour is
    // their alpha-weighted mean; front-to-back compositing into `image`.
    uint64_t render_pixels_multi(hipStream_t s, const RenderRequest& req, const std::vector<const NeRFRenderable*>& rs,
                                 uint32_t start, uint32_t n, float* image, const Settings& st);
};

struct NeRFRenderingController {
    explicit NeRFRenderingController(RenderPattern pattern = RenderPattern::LinearChunks, uint32_t batch_size = 0);
    void submit(std::shared_ptr<RenderRequest> request);
    void cancel();
    float min_step_size = NeRFConstants::min_step_size;
    uint32_t max_samples_per_ray = 0xFFFFFFFFu;
    uint64_t last_sample_count = 0;
    // render a single NeRF through the shared fixed-step loop too (tests)
    bool force_shared_steps = false;

private:
    RenderPattern pattern_;
    uint32_t batch_size_;
    Renderer renderer_;
    std::shared_ptr<RenderRequest> current_;
};

void register_renderer_bindings(pybind11::module_& m);

}  // namespace turbo
