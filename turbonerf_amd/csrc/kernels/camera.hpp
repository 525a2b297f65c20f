// World-space camera rays (reference Camera::global_ray_at_pixel_xy,
// src/models/camera.cuh:94-141), shared by the training and render ray
// kernels. Both include this after `#pragma clang fp contract(off)`: every
// product and sum is rounded on its own, in the order written, so the rays
// are bit-for-bit those of a plain C++ build of the same expressions.
#pragma once

#include "common.hpp"

namespace tn {

// Undistorted normalised point of the distorted (xd, yd) under the radial
// (k1 k2 k3) + tangential (p1 p2) lens model, by Newton's method from
// (xd, yd): at most 10 steps, stopping when the Jacobian's determinant falls
// below 1e-9 or both steps do (reference camera-kernels.cuh:12-105). The
// reference's C++ types are kept: terms with a double literal are evaluated
// in double and rounded to float once, the rest in float.
TN_DEVICE void lens_undistort(float xd, float yd, const float (&k)[5], float& xu, float& yu) {
    const float k1 = k[0], k2 = k[1], k3 = k[2], p1 = k[3], p2 = k[4];
    const float eps = 1e-9f;
    float x = xd, y = yd;
    for (int it = 0; it < 10; ++it) {
        // residual of the forward model at (x, y)
        const float r = x * x + y * y;
        const float d = (float)(1.0 + (double)(r * (k1 + r * (k2 + r * k3))));
        const float fx = d * x + 2.0f * p1 * x * y + p2 * (r + 2.0f * x * x) - xd;
        const float fy = d * y + 2.0f * p2 * x * y + p1 * (r + 2.0f * y * y) - yd;
        // its Jacobian
        const double rd = (double)r;
        const float d_r = (float)((double)k1 + rd * (2.0 * (double)k2 + rd * (3.0 * (double)k3)));
        const float d_x = (float)(2.0 * (double)x * (double)d_r);
        const float d_y = (float)(2.0 * (double)y * (double)d_r);
        const double p1d = (double)p1, p2d = (double)p2, xdb = (double)x, ydb = (double)y;
        const float fx_x = (float)((double)(d + d_x * x) + 2.0 * p1d * ydb + 6.0 * p2d * xdb);
        const float fx_y = (float)((double)(d_y * x) + 2.0 * p1d * xdb + 2.0 * p2d * ydb);
        const float fy_x = (float)((double)(d_x * y) + 2.0 * p2d * ydb + 2.0 * p1d * xdb);
        const float fy_y = (float)((double)(d + d_y * y) + 2.0 * p2d * xdb + 6.0 * p1d * ydb);
        const float det = fx_y * fy_x - fx_x * fy_y;
        if (fabsf(det) < eps) break;
        const float sx = (fx * fy_y - fy * fx_y) / det;
        const float sy = (fy * fx_x - fx * fy_x) / det;
        x += sx;
        y += sy;
        if (fabsf(sx) < eps && fabsf(sy) < eps) break;
    }
    xu = x;
    yu = y;
}

// Ray through pixel (x, y) in world space: origin on the near plane, unit
// direction. Pixels of a camera with any non-zero distortion parameter are
// undistorted first, in units of the focal length about the principal point.
TN_DEVICE void camera_ray(const tn_camera_t& cam, float x, float y, float& ox, float& oy, float& oz, float& dx,
                          float& dy, float& dz) {
    if (cam.dist_params[0] != 0.0f || cam.dist_params[1] != 0.0f || cam.dist_params[2] != 0.0f ||
        cam.dist_params[3] != 0.0f || cam.dist_params[4] != 0.0f) {
        const float xd = (x - cam.principal_point[0]) / cam.focal_length[0];
        const float yd = (y - cam.principal_point[1]) / cam.focal_length[1];
        float xu, yu;
        lens_undistort(xd, yd, cam.dist_params, xu, yu);
        x = xu * cam.focal_length[0] + cam.principal_point[0];
        y = yu * cam.focal_length[1] + cam.principal_point[1];
    }
    const float vx = (x + cam.offset_px[0] + 0.5f) / cam.focal_length[0];
    const float vy = (y + cam.offset_px[1] + 0.5f) / cam.focal_length[1];
    const float vz = 1.0f;
    const float v_len = sqrtf(vx * vx + vy * vy + vz * vz);
    const float* m = cam.transform;
    float gx = m[0] * vx + m[1] * vy + m[2] * vz;
    float gy = m[4] * vx + m[5] * vy + m[6] * vz;
    float gz = m[8] * vx + m[9] * vy + m[10] * vz;
    const float gl = sqrtf(gx * gx + gy * gy + gz * gz);
    gx = gx / gl;
    gy = gy / gl;
    gz = gz / gl;
    const float sx = v_len * gx, sy = v_len * gy, sz = v_len * gz;
    ox = m[3] + cam.near_plane * sx;
    oy = m[7] + cam.near_plane * sy;
    oz = m[11] + cam.near_plane * sz;
    dx = gx;
    dy = gy;
    dz = gz;
}

}  // namespace tn
