// File-backed datasets: transforms.json (the instant-ngp / nerfstudio layout)
// and PNG training images (reference src/models/dataset.cu:24-225).
#include <algorithm>
#include <charconv>
#include <cstdio>
#include <atomic>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <map>
#include <thread>
#include <utility>

#include "image_io.hpp"
#include "json.hpp"
#include "models.hpp"

namespace turbo {

namespace fs = std::filesystem;

namespace {

Transform4f matrix_rows(const JsonValue& m) {
    if (!m.is_array() || m.size() < 3) throw std::runtime_error("transform_matrix must have at least 3 rows");
    float v[12];
    for (int r = 0; r < 3; ++r) {
        const JsonValue& row = m.at((size_t)r);
        if (!row.is_array() || row.size() < 4) throw std::runtime_error("transform_matrix rows need 4 entries");
        for (int c = 0; c < 4; ++c) v[4 * r + c] = row.at((size_t)c).as_float("transform_matrix entry");
    }
    return Transform4f::from_array(v);
}

// Scale(s) * M: every entry of the 3x4 rows times s (translation included)
Transform4f scaled(const Transform4f& m, float s) {
    std::array<float, 12> a = m.to_array();
    for (float& x : a) x = s * x;
    return Transform4f::from_array(a.data());
}

}  // namespace

void Dataset::load_transforms() {
    if (!file_path.has_value()) throw std::runtime_error("No file path specified");
    std::ifstream in(*file_path, std::ios::binary);
    if (!in) throw std::runtime_error("Could not open file: " + *file_path);
    const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const JsonValue j = parse_json(text);
    const JsonValue& frames = j.at("frames");
    if (!frames.is_array()) throw std::runtime_error("transforms: \"frames\" must be an array");

    // scene-wide values and their defaults (dataset.cu:44-82)
    const float w = j.value("w", 0.0f);
    const float h = j.value("h", 0.0f);
    const int2v global_dims{(int)w, (int)h};
    const int aabb_size = std::min(j.value("aabb_scale", 16), 128);
    const float scene_scale = j.value("scene_scale", 1.0f);
    const float global_cx = j.value("cx", 0.5f * w);
    const float global_cy = j.value("cy", 0.5f * h);
    float global_fl_x, global_fl_y;
    if (j.contains("camera_angle_x")) {
        const float ca_x = j.at("camera_angle_x").as_float("camera_angle_x");
        const float ca_y = j.value("camera_angle_y", ca_x);
        global_fl_x = 0.5f * w / tanf(0.5f * ca_x);
        global_fl_y = 0.5f * h / tanf(0.5f * ca_y);
    } else {
        global_fl_x = j.value("fl_x", 1000.0f);
        global_fl_y = j.value("fl_y", global_fl_x);
    }
    const float global_near = j.value("near", 0.05f);
    const float global_far = j.value("far", 128.0f);
    const DistortionParams global_dist{j.value("k1", 0.0f), j.value("k2", 0.0f), j.value("k3", 0.0f),
                                       j.value("p1", 0.0f), j.value("p2", 0.0f)};

    std::vector<Camera> cams;
    std::vector<std::string> paths;
    const fs::path base_dir = fs::path(*file_path).parent_path();
    for (const JsonValue& frame : frames.array) {
        const fs::path image_path = base_dir / frame.at("file_path").as_string("file_path");
        if (!fs::exists(image_path)) continue;  // frames without an image file are skipped
        int2v dims{frame.value("w", global_dims.x), frame.value("h", global_dims.y)};
        if (dims.x <= 0 || dims.y <= 0) {
            // no size in the file: take it from the image's header
            if (!read_image_size(image_path.string(), dims.x, dims.y))
                throw std::runtime_error("transforms: no image size given and " + image_path.string() +
                                         " is not a PNG or JPEG");
        }
        const float near = scene_scale * frame.value("near", global_near);
        const float far = scene_scale * frame.value("far", global_far);
        const Transform4f camera_matrix = scaled(matrix_rows(frame.at("transform_matrix")).from_nerf(), scene_scale);
        const float cx = frame.value("cx", global_cx);
        const float cy = frame.value("cy", global_cy);
        const float fl_x = frame.value("fl_x", global_fl_x);
        const float fl_y = frame.value("fl_y", global_fl_y);
        const DistortionParams dist{frame.value("k1", global_dist.k1), frame.value("k2", global_dist.k2),
                                    frame.value("k3", global_dist.k3), frame.value("p1", global_dist.p1),
                                    frame.value("p2", global_dist.p2)};
        cams.emplace_back(dims, near, far, float2v{fl_x, fl_y}, float2v{cx, cy}, float2v{0.0f, 0.0f}, camera_matrix,
                          dist);
        paths.push_back(image_path.string());
    }
    if (cams.empty()) throw std::runtime_error("No valid images in this dataset!");

    // one image size per dataset: keep the most common one; among equally
    // common sizes the smallest (width, then height) wins (dataset.cu:136-176)
    std::map<std::pair<int, int>, int> counts;
    for (const Camera& c : cams) counts[{c.resolution.x, c.resolution.y}]++;
    contains_multiple_image_dims = counts.size() > 1;
    if (contains_multiple_image_dims) {
        std::pair<int, int> best{0, 0};
        int best_count = 0;
        for (const auto& kv : counts)
            if (kv.second > best_count) {
                best_count = kv.second;
                best = kv.first;
            }
        std::vector<Camera> kept;
        std::vector<std::string> kept_paths;
        for (size_t i = 0; i < cams.size(); ++i)
            if (cams[i].resolution.x == best.first && cams[i].resolution.y == best.second) {
                kept.push_back(cams[i]);
                kept_paths.push_back(paths[i]);
            }
        cams.swap(kept);
        paths.swap(kept_paths);
    }
    cameras = std::move(cams);
    image_paths = std::move(paths);
    bounding_box = BoundingBox((float)aabb_size);
    image_dimensions = cameras[0].resolution;
    n_pixels_per_image = (uint32_t)image_dimensions.x * (uint32_t)image_dimensions.y;
    unload_images();
}

namespace {
std::string num(float v) {
    if (!std::isfinite(v)) throw std::runtime_error("Dataset::to_json: a camera holds a non-finite value");
    char b[32];
    // 9 significant digits read back as the same float; to_chars ignores the locale
    const auto r = std::to_chars(b, b + sizeof(b), v, std::chars_format::general, 9);
    return std::string(b, r.ptr);
}
std::string quoted(const std::string& s) {
    std::string o = "\"";
    for (const char c : s) {
        if (c == '"' || c == '\\') o += '\\';
        if ((unsigned char)c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof(b), "\\u%04x", (unsigned)(unsigned char)c);
            o += b;
            continue;
        }
        o += c;
    }
    return o + "\"";
}
}  // namespace

std::string Dataset::to_json() const {
    if (!file_path.has_value() || image_paths.size() != cameras.size())
        throw std::runtime_error("Dataset::to_json: only datasets read from a transforms.json can be written");
    const fs::path base = fs::path(*file_path).parent_path();
    std::string o = "{\n  \"w\": " + std::to_string(image_dimensions.x) + ",\n  \"h\": " +
                    std::to_string(image_dimensions.y) + ",\n  \"aabb_scale\": " + num(bounding_box.size()) +
                    ",\n  \"scene_scale\": 1.0,\n  \"frames\": [";
    for (size_t i = 0; i < cameras.size(); ++i) {
        const Camera& c = cameras[i];
        const std::array<float, 12> m = c.transform.to_nerf().to_array();
        std::string mat = "[";
        for (int r = 0; r < 3; ++r)
            mat += "[" + num(m[4 * r]) + ", " + num(m[4 * r + 1]) + ", " + num(m[4 * r + 2]) + ", " +
                   num(m[4 * r + 3]) + "], ";
        mat += "[0.0, 0.0, 0.0, 1.0]]";
        const std::string rel = fs::path(image_paths[i]).lexically_relative(base).generic_string();
        o += std::string(i ? "," : "") + "\n    {\"near\": " + num(c.near) + ", \"far\": " + num(c.far) +
             ", \"cx\": " + num(c.principal_point.x) + ", \"cy\": " + num(c.principal_point.y) +
             ", \"fl_x\": " + num(c.focal_length.x) + ", \"fl_y\": " + num(c.focal_length.y) +
             ", \"k1\": " + num(c.dist_params.k1) + ", \"k2\": " + num(c.dist_params.k2) +
             ", \"k3\": " + num(c.dist_params.k3) + ", \"p1\": " + num(c.dist_params.p1) +
             ", \"p2\": " + num(c.dist_params.p2) + ", \"transform_matrix\": " + mat +
             ", \"file_path\": " + quoted(rel) + "}";
    }
    return o + "\n  ]\n}\n";
}

void Dataset::load_images() {
    if (image_paths.size() != cameras.size())
        throw std::runtime_error("Dataset::load_images: the dataset has no image files (in-memory dataset)");
    const size_t n = image_paths.size();
    const size_t bytes = (size_t)n_pixels_per_image * 4;
    std::vector<uint8_t> data(n * bytes);
    std::vector<std::string> errors(n);
    std::atomic<size_t> next{0};
    const unsigned n_threads = std::max(1u, std::min({std::thread::hardware_concurrency(), 16u, (unsigned)n}));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < n_threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) {
                try {
                    const Rgba8Image img = load_image_rgba8(image_paths[i]);
                    if (img.width != image_dimensions.x || img.height != image_dimensions.y)
                        throw std::runtime_error(image_paths[i] + " is " + std::to_string(img.width) + "x" +
                                                 std::to_string(img.height) + ", the dataset declares " +
                                                 std::to_string(image_dimensions.x) + "x" +
                                                 std::to_string(image_dimensions.y));
                    std::copy(img.rgba.begin(), img.rgba.end(), data.begin() + (std::ptrdiff_t)(i * bytes));
                } catch (const std::exception& e) {
                    errors[i] = e.what();
                }
            }
        });
    for (auto& th : pool) th.join();
    for (const std::string& e : errors)
        if (!e.empty()) throw std::runtime_error("Dataset::load_images: " + e);
    image_data.swap(data);
}

}  // namespace turbo
