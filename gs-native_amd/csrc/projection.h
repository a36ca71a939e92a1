// Projection of one Gaussian: view-space position, 2D covariance and conic, tile rectangle,
// depth, opacity and the SH basis of the view direction (reference
// src/core/render/renderer.cu:108-221), shared by the projection kernel and the per-Gaussian pass
// that projects the updated parameters for the next view. Every operation runs with FP
// contraction off and the shared FP32 transcendental functions, so both kernels write the same
// bits, and the integer outputs equal the host restatement bit for bit.
// Difference from the reference: when the 2D covariance is near-degenerate (det < 1e-3 * a * c,
// where the FP32 difference a*c - b*b loses more than ten bits), det and the conic are formed in
// FP64 from the FP32 a, b, c (whose products are exact in FP64).
#pragma once

#include "gsn_internal.h"

namespace gsn {

constexpr float kNearDegenerate = 1e-3f;

// Geometry of Gaussian i from its parameter rows 0-9 (row(r) returns row r). The caller has
// written record_counts[i] = 0; a culled Gaussian returns false and writes nothing else (its other
// outputs keep their previous contents, as in the reference). Otherwise writes depth, rectangle,
// record count, mean and conic + opacity, fills basis[0 .. (Degree+1)^2) and returns true.
template<int Degree, typename Row>
__device__ __forceinline__ bool project_geometry(Row row, const Camera& camera, int i, float2* means,
    float4* conic_opacities, float* depths, int4* tile_rects, int* record_counts, float (&basis)[kShCoeffCount]) {
#pragma clang fp contract(off)
    constexpr int sh_degree = Degree;
    float px0 = row(kPositionIndex + 0);
    float py0 = row(kPositionIndex + 1);
    float pz0 = row(kPositionIndex + 2);
    const float* Rc = camera.R_c_w;
    const float* tc = camera.t_c_w;
    float px = Rc[0] * px0 + Rc[1] * py0 + Rc[2] * pz0 + tc[0];
    float py = Rc[3] * px0 + Rc[4] * py0 + Rc[5] * pz0 + tc[1];
    float z = Rc[6] * px0 + Rc[7] * py0 + Rc[8] * pz0 + tc[2];
    float near_depth = fmaxf(camera.near_depth, kMinNearDepth);
    if (z <= near_depth || z > fmaxf(camera.far_depth, near_depth)) {
        return false;
    }
    float Rg[9];
    rotation_from_axis_angle(row(kRotationIndex + 0),
        row(kRotationIndex + 1), row(kRotationIndex + 2), Rg);
    float scale[3] = {det_expf(row(kLogScaleIndex + 0)),
        det_expf(row(kLogScaleIndex + 1)),
        det_expf(row(kLogScaleIndex + 2))};
    float M[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            M[r * 3 + c] = (Rc[r * 3] * Rg[c] + Rc[r * 3 + 1] * Rg[3 + c] + Rc[r * 3 + 2] * Rg[6 + c]) * scale[c];
        }
    }
    float cov[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cov[r * 3 + c] = M[r * 3] * M[c * 3] + M[r * 3 + 1] * M[c * 3 + 1] + M[r * 3 + 2] * M[c * 3 + 2];
        }
    }
    float limit_x = kFovClampScale * 0.5f * static_cast<float>(camera.width) / camera.fx;
    float limit_y = kFovClampScale * 0.5f * static_cast<float>(camera.height) / camera.fy;
    float x = fminf(fmaxf(px / z, -limit_x), limit_x) * z;
    float y = fminf(fmaxf(py / z, -limit_y), limit_y) * z;
    float J[6] = {camera.fx / z, 0.0f, -camera.fx * x / (z * z), 0.0f, camera.fy / z, -camera.fy * y / (z * z)};
    float JS[6];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            JS[r * 3 + c] = J[r * 3] * cov[c] + J[r * 3 + 1] * cov[3 + c] + J[r * 3 + 2] * cov[6 + c];
        }
    }
    float a = JS[0] * J[0] + JS[1] * J[1] + JS[2] * J[2] + kLowPassVariance;
    float b = JS[0] * J[3] + JS[1] * J[4] + JS[2] * J[5];
    float c = JS[3] * J[3] + JS[4] * J[4] + JS[5] * J[5] + kLowPassVariance;
    float det = a * c - b * b;
    bool degenerate = det < kNearDegenerate * (a * c);
    double det64 = 0.0;
    if (degenerate) {
        det64 = static_cast<double>(a) * static_cast<double>(c) - static_cast<double>(b) * static_cast<double>(b);
        det = static_cast<float>(det64);
    }
    if (!(det > 0.0f)) {
        return false;
    }
    float middle = 0.5f * (a + c);
    float radius = ceilf(3.0f * sqrtf(middle + sqrtf(fmaxf(0.1f, middle * middle - det))));
    float u = camera.fx * px / z + camera.cx;
    float v = camera.fy * py / z + camera.cy;
    if (!isfinite(radius) || !isfinite(u) || !isfinite(v)) {
        return false;
    }
    int4 rect;
    rect.x = static_cast<int>(fminf(static_cast<float>(camera.tile_count_x), fmaxf(0.0f, (u - radius) / kTileSize)));
    rect.y = static_cast<int>(fminf(static_cast<float>(camera.tile_count_y), fmaxf(0.0f, (v - radius) / kTileSize)));
    rect.z = static_cast<int>(
        fminf(static_cast<float>(camera.tile_count_x), fmaxf(0.0f, (u + radius + kTileSize - 1) / kTileSize)));
    rect.w = static_cast<int>(
        fminf(static_cast<float>(camera.tile_count_y), fmaxf(0.0f, (v + radius + kTileSize - 1) / kTileSize)));
    if (rect.x >= rect.z || rect.y >= rect.w) {
        return false;
    }
    depths[i] = z;
    tile_rects[i] = rect;
    record_counts[i] = (rect.z - rect.x) * (rect.w - rect.y);
    float4 conic;
    if (degenerate) {
        conic.x = static_cast<float>(static_cast<double>(c) / det64);
        conic.y = static_cast<float>(-static_cast<double>(b) / det64);
        conic.z = static_cast<float>(static_cast<double>(a) / det64);
    }
    else {
        float det_inv = 1.0f / det;
        conic.x = c * det_inv;
        conic.y = -b * det_inv;
        conic.z = a * det_inv;
    }
    conic.w = 1.0f / (1.0f + det_expf(-row(kOpacityIndex)));
    means[i] = make_float2(u, v);
    conic_opacities[i] = conic;

    // View direction from the camera centre (p + R^T t), normalised.
    float dx = px0 + Rc[0] * tc[0] + Rc[3] * tc[1] + Rc[6] * tc[2];
    float dy = py0 + Rc[1] * tc[0] + Rc[4] * tc[1] + Rc[7] * tc[2];
    float dz = pz0 + Rc[2] * tc[0] + Rc[5] * tc[1] + Rc[8] * tc[2];
    float length_inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
    float xs = dx * length_inv;
    float ys = dy * length_inv;
    float zs = dz * length_inv;
    basis[0] = kShC0;
    if (sh_degree >= 1) {
        basis[1] = -kShC1 * ys;
        basis[2] = kShC1 * zs;
        basis[3] = -kShC1 * xs;
    }
    if (sh_degree >= 2) {
        float xx = xs * xs, yy = ys * ys, zz = zs * zs;
        float xy = xs * ys, yz = ys * zs, xz = xs * zs;
        basis[4] = kShC2_0 * xy;
        basis[5] = kShC2_1 * yz;
        basis[6] = kShC2_2 * (2.0f * zz - xx - yy);
        basis[7] = kShC2_3 * xz;
        basis[8] = kShC2_4 * (xx - yy);
        if (sh_degree >= 3) {
            basis[9] = kShC3_0 * ys * (3.0f * xx - yy);
            basis[10] = kShC3_1 * xy * zs;
            basis[11] = kShC3_2 * ys * (4.0f * zz - xx - yy);
            basis[12] = kShC3_3 * zs * (2.0f * zz - 3.0f * xx - 3.0f * yy);
            basis[13] = kShC3_4 * xs * (4.0f * zz - xx - yy);
            basis[14] = kShC3_5 * zs * (xx - yy);
            basis[15] = kShC3_6 * xs * (xx - 3.0f * yy);
        }
    }
    return true;
}

// Adds one SH coefficient's term to a colour channel, in coefficient order (value starts at 0).
__device__ __forceinline__ float add_sh_term(float value, float basis, float coefficient) {
#pragma clang fp contract(off)
    return value + basis * coefficient;
}

// The colour channel written for a visible Gaussian.
__device__ __forceinline__ float finish_colour(float value) {
#pragma clang fp contract(off)
    return fmaxf(value + 0.5f, 0.0f);
}

} // namespace gsn
