// Renderer and NeRFRenderingController (see renderer.hpp).
//
// Each pass takes a chunk of pixel rays through the same occupancy-skipping
// march as training (count -> scan -> generate, capped per ray), runs the
// network on the compacted samples only, and composites them front to back
// with early termination. The reference instead marches every ray a fixed
// number of steps per iteration, tests occupancy per sample and synchronises
// with the host to count and compact (src/core/renderer.cu:298-645); here a
// pass needs one host read, to learn how many rays fitted the sample budget.
#include "renderer.hpp"

#include "image_io.hpp"

#include <algorithm>
#include <cstring>

namespace turbo {

void RenderingWorkspace::enlarge(uint32_t rays, uint32_t samples, size_t e) {
    if (rays <= ray_capacity && samples <= sample_capacity && density.size() >= (size_t)samples * e) return;
    ray_capacity = rays;
    sample_capacity = samples;
    ray_ori.resize(3 * (size_t)rays);
    ray_dir.resize(3 * (size_t)rays);
    ray_t.resize(rays);
    ray_tmax.resize(rays);
    rgba.resize(4 * (size_t)rays);
    pixel.resize(rays);
    n_steps.resize(rays);
    offset.resize(rays);
    info.resize(4);
    alive.resize(rays);
    scan_ws.resize(std::max<size_t>(tn_scan_workspace_bytes(rays), 4));
    to_local.resize(12 * sizeof(float));
    s_img.resize(samples);
    s_t.resize(samples);
    s_pos.resize(3 * (size_t)samples);
    s_dir.resize(3 * (size_t)samples);
    s_dt.resize(samples);
    s_m.resize(samples);
    s_dtn.resize(samples);
    features.resize(32 * (size_t)samples * 4);  // fp32 in both precisions
    density.resize((size_t)samples * e);
    color.resize(3 * (size_t)samples * e);
}

uint64_t Renderer::render_pixels(hipStream_t s, const RenderRequest& req, const NeRFRenderable& r, uint32_t start,
                                 uint32_t n, float* image, const Settings& st) {
    NeRF& nerf = r.proxy->nerfs[0];
    NerfNetwork& net = nerf.network;
    const tn_stream_t ts = as_tn(s);
    const uint32_t R = ws.ray_capacity, S = ws.sample_capacity;
    const tn_camera_t cam = req.camera.c();
    const tn_bbox_t bbox = r.bounding_box.c();
    const float aabb = r.proxy->training_bbox.get().size();
    // the reference's renderer picks a sample's cascade from the step
    // normalised by the scene size (rendering-kernels.cu:445,447)
    tn_occupancy_grid_t grid = nerf.occupancy_grid.view();
    grid.level_dt_scale = 1.0f / aabb;
    const float dt_min = st.min_step_size;
    const float dt_max = 8.0f * dt_min;  // reference renderer.cu:305
    const Transform4f& tr = r.proxy->transform.get();
    const float* to_local = nullptr;
    if (!(tr == Transform4f::Identity())) {
        const auto a = tr.inverse().to_array();
        TN_HIP(hipMemcpyAsync(ws.to_local.data(), a.data(), sizeof(a), hipMemcpyHostToDevice, s));
        to_local = reinterpret_cast<const float*>(ws.to_local.data());
    }
    uint64_t n_evals = 0;
    uint32_t p = start;
    const uint32_t end = start + n;
    uint32_t info_h[4];
    while (p < end) {
        if (req.is_canceled()) break;
        // a pass marches only about as many rays as the sample budget takes
        // (the rays past the kept prefix would be marched again next pass)
        uint32_t m = std::min(end - p, R);
        if (samples_per_ray > 0.0)
            m = std::min(m, (uint32_t)std::clamp(1.05 * (double)S / samples_per_ray, 256.0, (double)R));
        TN_TRY(tn_render_rays(m, p, R, cam, bbox, to_local, ws.ray_ori.data(), ws.ray_dir.data(), ws.ray_t.data(),
                              ws.ray_tmax.data(), ws.pixel.data(), ws.alive.data(), ts));
        TN_TRY(tn_march_count(m, nullptr, R, bbox, grid, NeRFConstants::cone_angle, dt_min, dt_max,
                              st.max_samples_per_ray, ws.ray_dir.data(), ws.ray_tmax.data(), ws.alive.data(),
                              ws.ray_ori.data(), ws.ray_t.data(), ws.n_steps.data(), ts));
        TN_TRY(tn_scan_select(1, m, nullptr, ws.n_steps.data(), S, ws.offset.data(), ws.info.data(), ws.scan_ws.data(),
                              ws.scan_ws.size(), ts));
        TN_TRY(tn_march_generate(m, ws.info.data(), R, S, bbox, 1.0f / aabb, grid, dt_min, dt_max,
                                 NeRFConstants::cone_angle, ws.pixel.data(), ws.ray_ori.data(), ws.ray_dir.data(),
                                 ws.ray_t.data(), ws.offset.data(), ws.alive.data(), ws.n_steps.data(), ws.s_img.data(),
                                 ws.s_t.data(), ws.s_pos.data(), ws.s_dir.data(), ws.s_dt.data(), ws.s_m.data(),
                                 ws.s_dtn.data(), ts));
        if (nerf.analytic_blobs.size()) {
            // synthetic scene: sigma and colour of the blob field
            TN_TRY(tn_blob_eval(ws.info.data(), S, S, aabb, nerf.analytic_blobs.data(),
                                (uint32_t)(nerf.analytic_blobs.size() / 8), ws.s_pos.data(),
                                reinterpret_cast<float*>(ws.density.data()), reinterpret_cast<float*>(ws.color.data()),
                                ts));
            TN_TRY(tn_composite_forward(0, ws.info.data(), S, S, R, st.min_transmittance, 0, ws.n_steps.data(),
                                        ws.offset.data(), ws.density.data(), ws.color.data(), ws.s_dt.data(),
                                        ws.rgba.data(), ts));
        } else {
            net.inference(s, nerf.params, S, ws.info.data(), S, ws.s_pos.data(), ws.s_dir.data(), ws.features.data(),
                          ws.density.data(), ws.color.data(), true);
            TN_TRY(tn_composite_forward(net.config.precision, ws.info.data(), S, S, R, st.min_transmittance, 1,
                                        ws.n_steps.data(), ws.offset.data(), ws.density.data(), ws.color.data(),
                                        ws.s_dt.data(), ws.rgba.data(), ts));
        }
        TN_TRY(tn_render_write_rgba(m, ws.info.data(), R, ws.pixel.data(), ws.alive.data(), ws.rgba.data(), image, ts));
        TN_HIP(hipMemcpyAsync(info_h, ws.info.data(), sizeof(info_h), hipMemcpyDeviceToHost, s));
        TN_HIP(hipStreamSynchronize(s));
        if (info_h[2] != 0 || info_h[0] == 0) {
            // mode 1 only fails when a single ray exceeds the sample budget
            if (info_h[0] == 0) throw std::runtime_error("render: a ray exceeds the per-pass sample budget");
        }
        p += info_h[0];
        n_evals += info_h[1];
        samples_per_ray = std::max(1.0, (double)info_h[1] / (double)info_h[0]);
    }
    return n_evals;
}

BoundingBox RepeaterEffect::get_bbox(const BoundingBox& b) const {
    BoundingBox o;
    o.min_x = std::min(extend_bbox.min_x, b.min_x);
    o.min_y = std::min(extend_bbox.min_y, b.min_y);
    o.min_z = std::min(extend_bbox.min_z, b.min_z);
    o.max_x = std::max(extend_bbox.max_x, b.max_x);
    o.max_y = std::max(extend_bbox.max_y, b.max_y);
    o.max_z = std::max(extend_bbox.max_z, b.max_z);
    return o;
}

void RenderingWorkspace::enlarge_multi(uint32_t rays, uint32_t samples, uint32_t n_nerfs) {
    if (m_trans.size() < rays) m_trans.resize(rays);
    if (m_dt.size() < samples) {
        m_pos.resize(3 * (size_t)samples);
        m_dt.resize(samples);
        m_net_pos.resize(3 * (size_t)samples);
        m_net_dir.resize(3 * (size_t)samples);
        m_net_dt.resize(samples);
        m_c_pos.resize(3 * (size_t)samples);
        m_c_dir.resize(3 * (size_t)samples);
        m_rgba.resize(4 * (size_t)samples);
        m_valid.resize(samples);
        m_offs.resize(samples);
        m_count.resize(samples);
        m_scan_ws.resize(std::max<size_t>(tn_scan_workspace_bytes(samples), 4));
    }
    m_alive_n.resize(4);
    if (m_nerfs.size() < n_nerfs * sizeof(tn_render_nerf_t)) m_nerfs.resize(n_nerfs * sizeof(tn_render_nerf_t));
}

uint64_t Renderer::render_pixels_multi(hipStream_t s, const RenderRequest& req,
                                       const std::vector<const NeRFRenderable*>& rs, uint32_t start, uint32_t n,
                                       float* image, const Settings& st) {
    const tn_stream_t ts = as_tn(s);
    const uint32_t R = ws.ray_capacity, S = ws.sample_capacity;
    const uint32_t n_nerfs = (uint32_t)rs.size();
    ws.enlarge_multi(R, S, n_nerfs);
    // each NeRF's placement: world -> NeRF transform, its render box grown by
    // its effects, an optional repeater and its training box
    std::vector<tn_render_nerf_t> nf(n_nerfs);
    for (uint32_t k = 0; k < n_nerfs; ++k) {
        const NeRFRenderable& r = *rs[k];
        tn_render_nerf_t& d = nf[k];
        const auto a = r.proxy->transform.get().inverse().to_array();
        std::copy(a.begin(), a.end(), d.to_local);
        BoundingBox box = r.bounding_box;
        d.repeat = 0;
        for (const auto& fx : r.spatial_effects) {
            box = fx->get_bbox(box);
            if (auto* rep = dynamic_cast<const RepeaterEffect*>(fx.get())) {
                d.repeat = 1;
                d.repeat_source = rep->source_bbox.c();
            }
        }
        d.render_box = box.c();
        const BoundingBox tb = r.proxy->training_bbox.get();
        d.inv_aabb = 1.0f / tb.size();
        d.training_min[0] = tb.min_x;
        d.training_min[1] = tb.min_y;
        d.training_min[2] = tb.min_z;
        d.training_size = tb.size();
    }
    TN_HIP(hipMemcpyAsync(ws.m_nerfs.data(), nf.data(), n_nerfs * sizeof(tn_render_nerf_t), hipMemcpyHostToDevice, s));
    const auto* d_nerfs = reinterpret_cast<const tn_render_nerf_t*>(ws.m_nerfs.data());
    const float dt_min = st.min_step_size, dt_max = 8.0f * dt_min;  // reference renderer.cu:305
    const uint32_t max_steps = std::min<uint32_t>(st.max_samples_per_ray, 1u << 20);
    uint64_t n_evals = 0;
    for (uint32_t p = start; p < start + n; p += R) {
        if (req.is_canceled()) break;
        const uint32_t m = std::min(start + n - p, R);
        const uint32_t k = std::max(1u, std::min(S / m, 1024u));  // fixed steps per ray per pass
        TN_TRY(tn_render_rays_multi(m, p, R, req.camera.c(), n_nerfs, d_nerfs, ws.ray_ori.data(), ws.ray_dir.data(),
                                    ws.ray_t.data(), ws.ray_tmax.data(), ws.n_steps.data(), ws.pixel.data(),
                                    ws.alive.data(), ws.m_trans.data(), ts));
        while (true) {
            uint32_t alive_h = 0;
            TN_TRY(tn_count_nonzero(m, ws.alive.data(), ws.m_alive_n.data(), ts));
            TN_HIP(hipMemcpyAsync(&alive_h, ws.m_alive_n.data(), sizeof(alive_h), hipMemcpyDeviceToHost, s));
            TN_HIP(hipStreamSynchronize(s));
            if (alive_h == 0 || req.is_canceled()) break;
            const uint32_t nq = m * k;
            TN_TRY(tn_render_march_fixed(m, k, S, NeRFConstants::cone_angle, dt_min, dt_max, max_steps,
                                         ws.alive.data(), ws.ray_ori.data(), ws.ray_dir.data(), R, ws.ray_t.data(),
                                         ws.ray_tmax.data(), ws.n_steps.data(), ws.m_pos.data(), ws.m_dt.data(), ts));
            for (int c = 0; c < 4; ++c)
                TN_TRY(tn_fill_u32(nq, reinterpret_cast<uint32_t*>(ws.m_rgba.data()) + (size_t)c * S, 0u, ts));
            TN_TRY(tn_fill_u32(nq, ws.m_count.data(), 0u, ts));
            for (uint32_t j = 0; j < n_nerfs; ++j) {
                NeRF& nerf = rs[j]->proxy->nerfs[0];
                tn_occupancy_grid_t grid = nerf.occupancy_grid.view();
                TN_TRY(tn_render_localize(m, k, S, nf[j], grid, ws.ray_dir.data(), R, ws.m_pos.data(), ws.m_dt.data(),
                                          ws.m_valid.data(), ws.m_net_pos.data(), ws.m_net_dir.data(),
                                          ws.m_net_dt.data(), ts));
                // the valid samples packed to the front (info[1] = their number)
                TN_TRY(tn_scan_select(1, nq, nullptr, ws.m_valid.data(), S, ws.m_offs.data(), ws.info.data(),
                                      ws.m_scan_ws.data(), ws.m_scan_ws.size(), ts));
                TN_TRY(tn_render_compact(nq, S, ws.m_valid.data(), ws.m_offs.data(), ws.m_net_pos.data(),
                                         ws.m_net_dir.data(), ws.m_c_pos.data(), ws.m_c_dir.data(), ts));
                int dtype, raw;
                if (nerf.analytic_blobs.size()) {
                    TN_TRY(tn_blob_eval(ws.info.data(), S, S, rs[j]->proxy->training_bbox.get().size(),
                                        nerf.analytic_blobs.data(), (uint32_t)(nerf.analytic_blobs.size() / 8),
                                        ws.m_c_pos.data(), reinterpret_cast<float*>(ws.density.data()),
                                        reinterpret_cast<float*>(ws.color.data()), ts));
                    dtype = 0;
                    raw = 0;
                } else {
                    nerf.network.inference(s, nerf.params, S, ws.info.data(), S, ws.m_c_pos.data(), ws.m_c_dir.data(),
                                           ws.features.data(), ws.density.data(), ws.color.data(), true);
                    dtype = nerf.network.config.precision;
                    raw = 1;
                }
                TN_TRY(tn_render_accumulate(dtype, nq, S, ws.m_valid.data(), ws.m_offs.data(), ws.density.data(),
                                            ws.color.data(), raw, ws.m_net_dt.data(), ws.m_rgba.data(),
                                            ws.m_count.data(), ts));
            }
            TN_TRY(tn_render_composite_fixed(m, k, S, st.min_transmittance, ws.m_dt.data(), ws.m_rgba.data(),
                                             ws.m_count.data(), ws.pixel.data(), ws.alive.data(), ws.m_trans.data(),
                                             image, ts));
            n_evals += (uint64_t)alive_h * k * n_nerfs;
        }
    }
    return n_evals;
}

void CUDARenderBuffer::save_image(const std::string& file_path) {
    if (width <= 0 || height <= 0 || rgba.size() != (size_t)width * height * 4)
        throw std::runtime_error("save_image: the render buffer holds no image");
    CPURenderBuffer host;
    host.set_size(width, height);
    TN_HIP(hipMemcpy(host.get_rgba(), rgba.data(), rgba.bytes(), hipMemcpyDeviceToHost));
    host.save_image(file_path);
}

void CPURenderBuffer::save_image(const std::string& file_path) {
    if (width <= 0 || height <= 0 || rgba.size() != (size_t)width * height * 4)
        throw std::runtime_error("save_image: the render buffer holds no image");
    std::vector<uint8_t> px(rgba.size());
    for (size_t i = 0; i < rgba.size(); ++i) {
        const float v = rgba[i] * 255.0f;
        px[i] = v >= 255.0f ? (uint8_t)255 : v > 0.0f ? (uint8_t)v : (uint8_t)0;  // truncation, as a cast does
    }
    write_png_rgba8(file_path, width, height, px.data());
}

NeRFRenderingController::NeRFRenderingController(RenderPattern pattern, uint32_t batch_size)
    : pattern_(pattern), batch_size_(batch_size == 0 ? (1u << 20) : batch_size) {}

void NeRFRenderingController::cancel() {
    if (current_) current_->cancel();
}

void NeRFRenderingController::submit(std::shared_ptr<RenderRequest> req) {
    if (!req || !req->output) throw std::runtime_error("submit: request without output");
    current_ = req;
    const hipStream_t s = DeviceManager::stream();
    const int W = req->camera.resolution.x, H = req->camera.resolution.y;
    if (req->output->width != W || req->output->height != H) req->output->set_size(W, H);
    auto* cpu = dynamic_cast<CPURenderBuffer*>(req->output);
    auto* dev = dynamic_cast<CUDARenderBuffer*>(req->output);
    if (!cpu && !dev) throw std::runtime_error("submit: unsupported render target");
    const uint32_t n_pixels = (uint32_t)W * (uint32_t)H;
    // the visible NeRFs that can render: one without spatial effects takes
    // the occupancy-skipping renderer, several (or effects) the reference's
    // shared fixed-step loop
    std::vector<const NeRFRenderable*> rs;
    for (auto& x : req->renderables)
        if (x.proxy && x.proxy->can_render && x.proxy->is_visible) rs.push_back(&x);
    // a device target is rendered into in place; a host one through a device image
    DeviceBuffer<float> staging;
    if (!dev) staging.resize(4 * (size_t)n_pixels);
    float* image = dev ? dev->data() : staging.data();
    TN_HIP(hipMemsetAsync(image, 0, 4 * (size_t)n_pixels * sizeof(float), s));
    last_sample_count = 0;
    auto publish = [&]() {
        if (cpu)
            TN_HIP(hipMemcpyAsync(cpu->get_rgba(), image, 4 * (size_t)n_pixels * sizeof(float), hipMemcpyDeviceToHost,
                                  s));
        TN_HIP(hipStreamSynchronize(s));
    };
    if (!rs.empty()) {
        size_t e = 4;
        for (const auto* r : rs) {
            const NeRF& nerf = r->proxy->nerfs[0];
            e = std::max(e, nerf.analytic_blobs.size() ? (size_t)4 : nerf.network.elem_bytes());
        }
        const bool multi = rs.size() > 1 || !rs[0]->spatial_effects.empty() || force_shared_steps;
        // tasks of batch_size/8 rays (reference nerf-rendering-controller.cu:84)
        const uint32_t rays_per_task = std::max(1u, batch_size_ / 8);
        renderer_.ws.enlarge(rays_per_task, std::max(batch_size_ * 4u, 1u << 22), e);
        Renderer::Settings st;
        st.min_step_size = min_step_size;
        st.max_samples_per_ray = max_samples_per_ray;
        auto render_all = [&](const RenderRequest& rq, uint32_t npx, float* img, bool progress, bool cancelable) {
            for (uint32_t p = 0; p < npx; p += rays_per_task) {
                if (cancelable && req->is_canceled()) break;
                const uint32_t m = std::min(rays_per_task, npx - p);
                last_sample_count += multi ? renderer_.render_pixels_multi(s, rq, rs, p, m, img, st)
                                           : renderer_.render_pixels(s, rq, *rs[0], p, m, img, st);
                if (progress && req->on_progress) req->on_progress((float)(p + m) / (float)npx);
            }
        };
        const bool grid = pattern_ != RenderPattern::LinearChunks;
        const bool want_preview = (req->flags & RenderFlags::Preview) == RenderFlags::Preview;
        const bool want_final = (req->flags & RenderFlags::Final) == RenderFlags::Final;
        const uint32_t n_tasks = (n_pixels + rays_per_task - 1) / rays_per_task + (grid ? 1 : 0);
        if (grid && want_preview) {
            // a preview of about 2^14 rays spread over the image (reference
            // nerf-rendering-controller.cu:85,111-116; the preview task is
            // not cancelable), each filling its block of pixels; rendered
            // with the camera scaled to the grid
            const double aspect = (double)W / (double)H;
            const uint32_t gh = std::max(1u, std::min((uint32_t)H, (uint32_t)std::lround(std::sqrt(16384.0 / aspect))));
            const uint32_t gw = std::max(1u, std::min((uint32_t)W, 16384u / gh));
            Camera pc = req->camera;
            pc.resolution = int2v{(int)gw, (int)gh};
            pc.focal_length = float2v{pc.focal_length.x * (float)gw / (float)W, pc.focal_length.y * (float)gh / (float)H};
            pc.principal_point =
                float2v{pc.principal_point.x * (float)gw / (float)W, pc.principal_point.y * (float)gh / (float)H};
            RenderRequest prq(pc, req->renderables, nullptr, req->modifiers, RenderFlags::Preview);
            DeviceBuffer<float> pimg(4 * (size_t)gw * gh);
            pimg.zero(s);
            render_all(prq, gw * gh, pimg.data(), false, false);
            TN_TRY(tn_render_expand_nearest(gw, gh, (uint32_t)W, (uint32_t)H, pimg.data(), image, as_tn(s)));
            publish();
            if (req->on_progress) req->on_progress(1.0f / (float)n_tasks);
        }
        if (want_final) {
            TN_HIP(hipMemsetAsync(image, 0, 4 * (size_t)n_pixels * sizeof(float), s));
            render_all(*req, n_pixels, image, true, true);
        }
    }
    publish();
    if (req->is_canceled()) {
        if (req->on_cancel) req->on_cancel();
    } else if (req->on_complete) {
        req->on_complete();
    }
    current_.reset();
}

}  // namespace turbo
