// Projection + SH colour + tile rectangle (one thread per Gaussian, coalesced SoA row reads).
// Follows reference src/core/render/renderer.cu:108-221; the per-Gaussian math is in
// projection.h.
#include "projection.h"

namespace gsn {
namespace {

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
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        record_counts[count] = 0;
    }
    if (i >= count) {
        return;
    }
    record_counts[i] = 0;
    const int64_t n = count;
    float basis[kShCoeffCount];
    if (!project_geometry<Degree>([&](int r) { return stream_load<Streaming>(params + r * n + i); }, camera, i, means,
            conic_opacities, depths, tile_rects, record_counts, basis)) {
        return;
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
            value = add_sh_term(value, basis[j], coefficients[ch][j]);
        }
        colors[static_cast<int64_t>(i) * 3 + ch] = finish_colour(value);
    }
}

} // namespace

int record_projection(gsn_context* context, int count, float* means, float* conic_opacities, float* colors,
    float* depths, int* tile_rects, int* record_offsets) {
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
    return GSN_OK;
}

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
    GSN_TRY(record_projection(context, count, means, conic_opacities, colors, depths, tile_rects, record_offsets));
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
