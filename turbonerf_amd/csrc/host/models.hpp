// Host data model mirroring the reference's device-copyable PODs
// (src/models/*, src/math/transform4f.cuh). Each converts to the plain-C
// struct the kernel library takes (tn_types.h).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <optional>
#include <string>
#include <vector>

#include "common.hpp"

namespace turbo {

struct int2v { int x = 0, y = 0; };
struct float2v { float x = 0.0f, y = 0.0f; };

// 3x4 affine transform, row-major (reference src/math/transform4f.cuh)
struct Transform4f {
    float m00 = 1, m01 = 0, m02 = 0, m03 = 0;
    float m10 = 0, m11 = 1, m12 = 0, m13 = 0;
    float m20 = 0, m21 = 0, m22 = 1, m23 = 0;

    static Transform4f Identity() { return Transform4f{}; }
    static Transform4f from_array(const float* p) {
        Transform4f t;
        float* d[12] = {&t.m00, &t.m01, &t.m02, &t.m03, &t.m10, &t.m11, &t.m12, &t.m13, &t.m20, &t.m21, &t.m22, &t.m23};
        for (int k = 0; k < 12; ++k) *d[k] = p[k];
        return t;
    }
    std::array<float, 12> to_array() const {
        return {m00, m01, m02, m03, m10, m11, m12, m13, m20, m21, m22, m23};
    }
    // inverse of the affine map (reference transform4f.cuh:208-236)
    Transform4f inverse() const {
        const float a = m11 * m22 - m12 * m21, b = m10 * m22 - m12 * m20, c = m10 * m21 - m11 * m20;
        const float det = m00 * a - m01 * b + m02 * c;
        const float id = 1.0f / det;
        const float d = m12 * m23 - m13 * m22, e = m11 * m23 - m13 * m21, f = m10 * m23 - m13 * m20;
        Transform4f r;
        r.m00 = a * id;
        r.m01 = -(m01 * m22 - m02 * m21) * id;
        r.m02 = (m01 * m12 - m02 * m11) * id;
        r.m03 = -(m01 * d - m02 * e + m03 * a) * id;
        r.m10 = -b * id;
        r.m11 = (m00 * m22 - m02 * m20) * id;
        r.m12 = -(m00 * m12 - m02 * m10) * id;
        r.m13 = (m00 * d - m02 * f + m03 * b) * id;
        r.m20 = c * id;
        r.m21 = -(m00 * m21 - m01 * m20) * id;
        r.m22 = (m00 * m11 - m01 * m10) * id;
        r.m23 = -(m00 * e - m01 * f + m03 * c) * id;
        return r;
    }
    // NeRF (OpenGL-style) camera to this project's frame: flip the y and z
    // columns, then roll rows xyz -> yzx (reference transform4f.cuh:239-262)
    Transform4f from_nerf() const {
        Transform4f r = *this;
        r.m01 = -r.m01; r.m11 = -r.m11; r.m21 = -r.m21;
        r.m02 = -r.m02; r.m12 = -r.m12; r.m22 = -r.m22;
        const Transform4f t = r;
        r.m00 = t.m10; r.m01 = t.m11; r.m02 = t.m12; r.m03 = t.m13;
        r.m10 = t.m20; r.m11 = t.m21; r.m12 = t.m22; r.m13 = t.m23;
        r.m20 = t.m00; r.m21 = t.m01; r.m22 = t.m02; r.m23 = t.m03;
        return r;
    }
    Transform4f to_nerf() const {
        Transform4f t = *this;
        Transform4f r = t;
        // undo the roll (yzx -> xyz), then flip the y and z columns back
        r.m10 = t.m00; r.m11 = t.m01; r.m12 = t.m02; r.m13 = t.m03;
        r.m20 = t.m10; r.m21 = t.m11; r.m22 = t.m12; r.m23 = t.m13;
        r.m00 = t.m20; r.m01 = t.m21; r.m02 = t.m22; r.m03 = t.m23;
        r.m01 = -r.m01; r.m11 = -r.m11; r.m21 = -r.m21;
        r.m02 = -r.m02; r.m12 = -r.m12; r.m22 = -r.m22;
        return r;
    }
    bool operator==(const Transform4f& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

// reference src/models/bounding-box.cuh:8-154
struct BoundingBox {
    float min_x = 0, min_y = 0, min_z = 0, max_x = 0, max_y = 0, max_z = 0;
    BoundingBox() = default;
    explicit BoundingBox(float size)
        : min_x(-0.5f * size), min_y(-0.5f * size), min_z(-0.5f * size),
          max_x(0.5f * size), max_y(0.5f * size), max_z(0.5f * size) {}
    float size() const { return max_x - min_x; }
    tn_bbox_t c() const { return tn_bbox_t{min_x, min_y, min_z, max_x, max_y, max_z}; }
    bool operator==(const BoundingBox& o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct DistortionParams {
    float k1 = 0, k2 = 0, k3 = 0, p1 = 0, p2 = 0;
    bool is_any_nonzero() const { return k1 != 0 || k2 != 0 || k3 != 0 || p1 != 0 || p2 != 0; }
};

// reference src/models/camera.cuh:45-91
struct Camera {
    int2v resolution;
    float near = 0, far = 0;
    float2v focal_length, principal_point, shift;
    Transform4f transform;
    DistortionParams dist_params;
    bool show_image_planes = true;

    Camera() = default;
    Camera(int2v res, float near_, float far_, float2v focal, float2v pp, float2v shift_, Transform4f t,
           DistortionParams dp = DistortionParams())
        : resolution(res), near(near_), far(far_), focal_length(focal), principal_point(pp), shift(shift_),
          transform(t), dist_params(dp) {}

    tn_camera_t c() const {
        tn_camera_t o{};
        o.resolution[0] = resolution.x;
        o.resolution[1] = resolution.y;
        o.near_plane = near;
        o.far_plane = far;
        o.focal_length[0] = focal_length.x;
        o.focal_length[1] = focal_length.y;
        o.resolution_f[0] = (float)resolution.x;
        o.resolution_f[1] = (float)resolution.y;
        o.principal_point[0] = principal_point.x;
        o.principal_point[1] = principal_point.y;
        o.shift[0] = shift.x;
        o.shift[1] = shift.y;
        o.offset_px[0] = shift.x * (float)resolution.x - principal_point.x;
        o.offset_px[1] = shift.y * (float)resolution.y - principal_point.y;
        const auto a = transform.to_array();
        std::memcpy(o.transform, a.data(), sizeof(o.transform));
        o.dist_params[0] = dist_params.k1;
        o.dist_params[1] = dist_params.k2;
        o.dist_params[2] = dist_params.k3;
        o.dist_params[3] = dist_params.p1;
        o.dist_params[4] = dist_params.p2;
        o.show_image_planes = show_image_planes ? 1 : 0;
        return o;
    }
    bool operator==(const Camera& o) const {
        return resolution.x == o.resolution.x && resolution.y == o.resolution.y && near == o.near && far == o.far &&
               focal_length.x == o.focal_length.x && focal_length.y == o.focal_length.y &&
               principal_point.x == o.principal_point.x && principal_point.y == o.principal_point.y &&
               shift.x == o.shift.x && shift.y == o.shift.y && transform == o.transform;
    }
    bool operator!=(const Camera& o) const { return !(*this == o); }
};

// Training images and cameras (reference src/models/dataset.h). Images are
// RGBA8, all of image_dimensions, concatenated. Loading transforms.json and
// image files is outside this build; datasets are constructed in memory.
// An image file of a dataset and its size (reference src/models/training-image.cuh:13-23)
struct TrainingImage {
    std::string file_path;
    int2v dimensions;
};

// A training set: cameras and their RGBA8 images, either given in memory or
// described by a transforms.json file (reference src/models/dataset.h:19-43).
struct Dataset {
    std::vector<Camera> cameras;
    std::vector<uint8_t> image_data;        // n_images x H x W x RGBA8, once loaded
    std::vector<std::string> image_paths;   // per camera, for file-backed datasets
    int2v image_dimensions;
    uint32_t n_pixels_per_image = 0;
    BoundingBox bounding_box;
    std::optional<std::string> file_path;
    bool contains_multiple_image_dims = false;

    Dataset() = default;
    explicit Dataset(const std::string& path) : file_path(path) {}
    Dataset(const BoundingBox& bbox, const std::vector<Camera>& cams, const uint8_t* rgba, int width, int height)
        : cameras(cams), image_dimensions{width, height}, n_pixels_per_image((uint32_t)width * (uint32_t)height),
          bounding_box(bbox) {
        image_data.assign(rgba, rgba + (size_t)cams.size() * n_pixels_per_image * 4);
    }
    size_t n_images() const { return cameras.size(); }
    // Reads cameras, image paths and the bounding box from file_path
    // (reference src/models/dataset.cu:24-188).
    void load_transforms();
    // Decodes every image file into image_data, on several threads
    // (reference dataset.cu:202-225).
    void load_images();
    void unload_images() { std::vector<uint8_t>().swap(image_data); }
    bool is_loaded_cpu() const { return image_data.size() == (size_t)n_images() * n_pixels_per_image * 4; }
    std::vector<TrainingImage> images() const {
        std::vector<TrainingImage> out;
        for (const std::string& p : image_paths) out.push_back({p, image_dimensions});
        return out;
    }
    // The dataset as transforms.json text: NeRF-convention matrices, image
    // paths relative to file_path's folder, scene scale 1 (reference
    // dataset.cu:228-288).
    std::string to_json() const;
    Dataset copy() const { return *this; }
};

// An updatable property with a dirty flag (reference src/models/updatable-property.cuh)
template <typename T>
struct UpdatableProperty {
    T value{};
    bool dirty = true;
    const T& get() const { return value; }
    void set(const T& v) { value = v; dirty = true; }
    bool is_dirty() const { return dirty; }
    void set_dirty(bool d) { dirty = d; }
};

}  // namespace turbo
