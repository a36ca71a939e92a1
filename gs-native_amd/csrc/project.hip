// Projection + SH colour + tile rectangle (one thread per Gaussian, coalesced SoA row reads).
// Follows reference src/core/render/renderer.cu:108-221. Every operation that feeds the tile
// rectangle or the depth key runs with FP contraction off and the shared FP32 transcendental
// functions, so the integer outputs equal the host restatement bit for bit.
// Difference from the reference: when the 2D covariance is near-degenerate (det < 1e-3 * a * c,
// where the FP32 difference a*c - b*b loses more than ten bits), det and the conic are formed in
// FP64 from the FP32 a, b, c (whose products are exact in FP64).
#include "gsn_internal.h"

namespace gsn {
namespace {

constexpr float kNearDegenerate = 1e-3f;

// Parameter rows are read once per projection: when every row starts on a 128-byte line (count a
// multiple of 32, aligned buffer), non-temporal loads keep them from displacing data that later
// kernels of the step read (measured 0.22 -> 0.17 ms at 3M Gaussians, SH degree 3). Otherwise a
// wave's 256-byte row segment straddles three lines, and ordinary loads, which leave the shared
// end lines in L2 for the neighbouring wave, are faster.
template<bool Streaming>
__device__ __forceinline__ float stream_load(const float* address) {
    return Streaming ? __builtin_nontemporal_load(address) : *address;
}

// Degree is a template parameter so that the SH loops unroll and all coefficient loads of a
// Gaussian are in flight together (a runtime-bounded loop issued them one dependent step at a time).
template<int Degree, bool Streaming>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ params, Camera camera, float2* means,
    float4* conic_opacities, float* colors, float* depths, int4* tile_rects, int* record_counts, int count) {
#pragma clang fp contract(off)
    constexpr int sh_degree = Degree;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        record_counts[count] = 0;
    }
    if (i >= count) {
        return;
    }
    record_counts[i] = 0;
    const int64_t n = count;
    float px0 = stream_load<Streaming>(params + (kPositionIndex + 0) * n + i);
    float py0 = stream_load<Streaming>(params + (kPositionIndex + 1) * n + i);
    float pz0 = stream_load<Streaming>(params + (kPositionIndex + 2) * n + i);
    const float* Rc = camera.R_c_w;
    const float* tc = camera.t_c_w;
    float px = Rc[0] * px0 + Rc[1] * py0 + Rc[2] * pz0 + tc[0];
    float py = Rc[3] * px0 + Rc[4] * py0 + Rc[5] * pz0 + tc[1];
    float z = Rc[6] * px0 + Rc[7] * py0 + Rc[8] * pz0 + tc[2];
    float near_depth = fmaxf(camera.near_depth, kMinNearDepth);
    if (z <= near_depth || z > fmaxf(camera.far_depth, near_depth)) {
        return;
    }
    float Rg[9];
    rotation_from_axis_angle(stream_load<Streaming>(params + (kRotationIndex + 0) * n + i),
        stream_load<Streaming>(params + (kRotationIndex + 1) * n + i), stream_load<Streaming>(params + (kRotationIndex + 2) * n + i), Rg);
    float scale[3] = {det_expf(stream_load<Streaming>(params + (kLogScaleIndex + 0) * n + i)),
        det_expf(stream_load<Streaming>(params + (kLogScaleIndex + 1) * n + i)),
        det_expf(stream_load<Streaming>(params + (kLogScaleIndex + 2) * n + i))};
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
        return;
    }
    float middle = 0.5f * (a + c);
    float radius = ceilf(3.0f * sqrtf(middle + sqrtf(fmaxf(0.1f, middle * middle - det))));
    float u = camera.fx * px / z + camera.cx;
    float v = camera.fy * py / z + camera.cy;
    if (!isfinite(radius) || !isfinite(u) || !isfinite(v)) {
        return;
    }
    int4 rect;
    rect.x = static_cast<int>(fminf(static_cast<float>(camera.tile_count_x), fmaxf(0.0f, (u - radius) / kTileSize)));
    rect.y = static_cast<int>(fminf(static_cast<float>(camera.tile_count_y), fmaxf(0.0f, (v - radius) / kTileSize)));
    rect.z = static_cast<int>(
        fminf(static_cast<float>(camera.tile_count_x), fmaxf(0.0f, (u + radius + kTileSize - 1) / kTileSize)));
    rect.w = static_cast<int>(
        fminf(static_cast<float>(camera.tile_count_y), fmaxf(0.0f, (v + radius + kTileSize - 1) / kTileSize)));
    if (rect.x >= rect.z || rect.y >= rect.w) {
        return;
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
    conic.w = 1.0f / (1.0f + det_expf(-stream_load<Streaming>(params + kOpacityIndex * n + i)));
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
    float basis[kShCoeffCount];
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
    constexpr int coeffs = (Degree + 1) * (Degree + 1);
    float coefficients[3][coeffs];
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
#pragma unroll
        for (int j = 0; j < coeffs; ++j) {
            coefficients[ch][j] = stream_load<Streaming>(params + (kShIndex + ch * kShCoeffCount + j) * n + i);
        }
    }
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float value = 0.0f;
#pragma unroll
        for (int j = 0; j < coeffs; ++j) {
            value += basis[j] * coefficients[ch][j];
        }
        colors[static_cast<int64_t>(i) * 3 + ch] = fmaxf(value + 0.5f, 0.0f);
    }
}

} // namespace
} // namespace gsn

extern "C" int gsn_project_gaussians(gsn_context* context, const float* params, int count, const gsn_camera* camera,
    int sh_degree, float* means, float* conic_opacities, float* colors, float* depths, int* tile_rects,
    int* record_offsets) {
    using namespace gsn;
    if (!context || !camera || count < 0 || !record_offsets) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_project_gaussians: invalid arguments");
    }
    if (count > 0 && (!params || !means || !conic_opacities || !colors || !depths || !tile_rects)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_project_gaussians: null buffer");
    }
    // New projection outputs: any depth order queued for earlier ones is stale.
    context->ordered_depths = nullptr;
    context->ordered_offsets = nullptr;
    context->ordered_count = -1;
    // Queued backward preparation of the previous projection may still read its outputs.
    GSN_TRY(join_backward_prefetch(context));
    context->projected_depths = count > 0 ? depths : nullptr;
    context->projected_offsets = record_offsets;
    context->projected_count = count;
    context->projected_means = means;
    context->projected_conics = conic_opacities;
    context->projected_colors = colors;
    context->projected_rects = tile_rects;
    context->projected_record_offsets = record_offsets;
    context->projected_gaussians = count;
    ++context->projection_generation;
    if (count == 0) {
        return check_hip(hipMemsetAsync(record_offsets, 0, sizeof(int), context->stream), "hipMemsetAsync");
    }
    Camera cam;
    static_assert(sizeof(Camera) == sizeof(gsn_camera), "camera layout");
    memcpy(&cam, camera, sizeof(cam));
    StageScope scope(context, kStageProject);
    unsigned grid = grid_for(count, 256);
    float2* m = reinterpret_cast<float2*>(means);
    float4* q = reinterpret_cast<float4*>(conic_opacities);
    int4* r = reinterpret_cast<int4*>(tile_rects);
    const bool streaming = rows_line_aligned(params, count);
    switch (clamp_sh_degree(sh_degree)) {
    case 0:
        if (streaming) {
            project_kernel<0, true><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        else {
            project_kernel<0, false><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        break;
    case 1:
        if (streaming) {
            project_kernel<1, true><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        else {
            project_kernel<1, false><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        break;
    case 2:
        if (streaming) {
            project_kernel<2, true><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        else {
            project_kernel<2, false><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        break;
    default:
        if (streaming) {
            project_kernel<3, true><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        else {
            project_kernel<3, false><<<grid, 256, 0, context->stream>>>(params, cam, m, q, colors, depths, r,
                record_offsets, count);
        }
        break;
    }
    return check_launch("project_kernel");
}
