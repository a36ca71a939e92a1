// Inputs and training labels of the learned densification proposal (no counterpart in the
// reference; the quantities reuse the reference's density statistics, density_controller.cu:
// 83 signed gradient mean, 97 absolute gradient mean, observation counts).
//
// Between two density steps every observation of a Gaussian updates a small state, rows of a
// [6][stride] float buffer: 0 max |screen gradient|, 1 max absolute screen gradient, 2 depth sum,
// 3 depth min (0 = none yet), 4 depth max, 5 max footprint radius in pixels (three standard
// deviations along the major axis of the 2D covariance, recovered from the conic).
// At a density step the 16 features of each Gaussian are
//   0 mean gradient / threshold          1 mean absolute gradient / threshold
//   2 log1p(max gradient / threshold)    3 log1p(max absolute gradient / threshold)
//   4 log1p(observations)                5-7 log-scales (descending) - log(split threshold)
//   8 opacity                            9-11 log(depth mean, min, max / scene extent)
//   12-14 SH DC (R, G, B)                15 log1p(max footprint radius)
// with the observation-derived entries 0 for Gaussians never observed in the interval.
#include "gsn_internal.h"

#include <cmath>

namespace gsn {
namespace {

constexpr int kFeatureRows = 6;
constexpr int kFeatures = 16;

__global__ __launch_bounds__(256) void feature_accumulate_kernel(const float* __restrict__ statistics,
    const float* __restrict__ absolute_statistics, const float* __restrict__ depths, const float4* __restrict__ conics,
    int count, float* __restrict__ state, int64_t stride) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count || statistics[static_cast<int64_t>(count) + id] == 0.0f) {
        return;
    }
    float gradient = statistics[id];
    float absolute = absolute_statistics[id];
    float depth = depths[id];
    float4 conic = conics[id];
    float half_trace = 0.5f * (conic.x + conic.z);
    float half_difference = 0.5f * (conic.x - conic.z);
    float smallest = half_trace - sqrtf(half_difference * half_difference + conic.y * conic.y);
    float radius = smallest > 1e-8f ? 3.0f * rsqrtf(smallest) : 1e4f;
    radius = fminf(radius, 1e4f);
    state[id] = fmaxf(state[id], gradient);
    state[stride + id] = fmaxf(state[stride + id], absolute);
    state[2 * stride + id] += depth;
    float minimum = state[3 * stride + id];
    state[3 * stride + id] = minimum > 0.0f ? fminf(minimum, depth) : depth;
    state[4 * stride + id] = fmaxf(state[4 * stride + id], depth);
    state[5 * stride + id] = fmaxf(state[5 * stride + id], radius);
}

__global__ __launch_bounds__(256) void feature_build_kernel(const float* __restrict__ params, int count,
    const float* __restrict__ gradient_sums, const float* __restrict__ absolute_gradient_sums,
    const uint32_t* __restrict__ observation_counts, const float* __restrict__ state, int64_t stride,
    float inverse_threshold, float log_split_threshold, float inverse_extent, float* __restrict__ features) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    const int64_t n = count;
    float f[kFeatures];
    uint32_t observations = observation_counts[id];
    float inverse_observations = observations ? 1.0f / static_cast<float>(observations) : 0.0f;
    float mean = gradient_sums[id] * inverse_observations;
    float absolute_mean = absolute_gradient_sums[id] * inverse_observations;
    f[0] = mean == mean ? mean * inverse_threshold : 0.0f;
    f[1] = absolute_mean == absolute_mean ? absolute_mean * inverse_threshold : 0.0f;
    f[2] = observations ? log1pf(state[id] * inverse_threshold) : 0.0f;
    f[3] = observations ? log1pf(state[stride + id] * inverse_threshold) : 0.0f;
    f[4] = log1pf(static_cast<float>(observations));
    float s0 = params[(kLogScaleIndex + 0) * n + id];
    float s1 = params[(kLogScaleIndex + 1) * n + id];
    float s2 = params[(kLogScaleIndex + 2) * n + id];
    float high = fmaxf(s0, fmaxf(s1, s2));
    float low = fminf(s0, fminf(s1, s2));
    float middle = s0 + s1 + s2 - high - low;
    f[5] = high - log_split_threshold;
    f[6] = middle - log_split_threshold;
    f[7] = low - log_split_threshold;
    f[8] = 1.0f / (1.0f + expf(-params[kOpacityIndex * n + id]));
    if (observations) {
        f[9] = logf(state[2 * stride + id] * inverse_observations * inverse_extent);
        f[10] = logf(state[3 * stride + id] * inverse_extent);
        f[11] = logf(state[4 * stride + id] * inverse_extent);
        f[15] = log1pf(state[5 * stride + id]);
    }
    else {
        f[9] = f[10] = f[11] = f[15] = 0.0f;
    }
    f[12] = params[(kShIndex + 0) * n + id];
    f[13] = params[(kShIndex + kShCoeffCount) * n + id];
    f[14] = params[(kShIndex + 2 * kShCoeffCount) * n + id];
    float4* out = reinterpret_cast<float4*>(features + static_cast<int64_t>(id) * kFeatures);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        out[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
    }
}

__global__ __launch_bounds__(256) void feature_clear_kernel(float* state, int64_t total) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < total) {
        state[i] = 0.0f;
    }
}

__device__ __forceinline__ bool useful(int j, int count, const float* params, const uint32_t* observation_counts,
    float min_opacity_logit) {
    return j >= 0 && j < count && observation_counts[j] > 0 &&
        params[static_cast<int64_t>(kOpacityIndex) * count + j] >= min_opacity_logit;
}

// Outcome of the previous density step for each Gaussian that existed before it: -1 pruned,
// 0 kept (or densified without benefit), 1 cloned and both copies still useful, 2 split and both
// children still useful. Useful = observed since that step and opacity >= the given minimum.
__global__ __launch_bounds__(256) void outcome_labels_kernel(const Int4* __restrict__ decisions,
    const Int4* __restrict__ offsets, Int4 totals, int previous_count, const float* __restrict__ params,
    const uint32_t* __restrict__ observation_counts, int count, float min_opacity_logit, int* __restrict__ labels) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= previous_count) {
        return;
    }
    Int4 d = decisions[id];
    Int4 o = offsets[id];
    int label;
    if (!d.x && !d.y && !d.w) {
        label = -1;
    }
    else if (d.y) {
        bool both = useful(o.x, count, params, observation_counts, min_opacity_logit) &&
            useful(totals.x + o.y, count, params, observation_counts, min_opacity_logit);
        label = both ? 1 : 0;
    }
    else if (d.w) {
        int base = totals.x + totals.y + o.w;
        bool both = useful(base, count, params, observation_counts, min_opacity_logit) &&
            useful(base + totals.w, count, params, observation_counts, min_opacity_logit);
        label = both ? 2 : 0;
    }
    else {
        label = 0;
    }
    labels[id] = label;
}

} // namespace
} // namespace gsn

extern "C" int gsn_density_features_clear(gsn_context* context, float* state, int stride) {
    using namespace gsn;
    if (!context || stride < 0 || (stride > 0 && !state)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_features_clear: invalid arguments");
    }
    int64_t total = static_cast<int64_t>(kFeatureRows) * stride;
    if (total == 0) {
        return GSN_OK;
    }
    feature_clear_kernel<<<grid_for(total, 256), 256, 0, context->stream>>>(state, total);
    return check_launch("feature_clear_kernel");
}

extern "C" int gsn_density_features_accumulate(gsn_context* context, const float* statistics,
    const float* absolute_statistics, const float* depths, const float* conic_opacities, int count, float* state,
    int stride) {
    using namespace gsn;
    if (!context || count < 0 || stride < count) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_features_accumulate: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    if (!statistics || !absolute_statistics || !depths || !conic_opacities || !state) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_features_accumulate: null buffer");
    }
    StageScope scope(context, kStageAccumulate);
    feature_accumulate_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(statistics, absolute_statistics,
        depths, reinterpret_cast<const float4*>(conic_opacities), count, state, stride);
    return check_launch("feature_accumulate_kernel");
}

extern "C" int gsn_density_features(gsn_context* context, const float* params, int count,
    const float* gradient_sums, const float* absolute_gradient_sums, const uint32_t* observation_counts,
    const float* state, int stride, float gradient_threshold, float split_threshold, float scene_extent,
    float* features) {
    using namespace gsn;
    if (!context || count < 0 || stride < count || !(gradient_threshold > 0.0f) || !(split_threshold > 0.0f) ||
        !(scene_extent > 0.0f)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_features: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    if (!params || !gradient_sums || !absolute_gradient_sums || !observation_counts || !state || !features) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_features: null buffer");
    }
    StageScope scope(context, kStageDensity);
    feature_build_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(params, count, gradient_sums,
        absolute_gradient_sums, observation_counts, state, stride, 1.0f / gradient_threshold, logf(split_threshold),
        1.0f / scene_extent, features);
    return check_launch("feature_build_kernel");
}

extern "C" int gsn_density_outcomes(gsn_context* context, const int32_t* decisions, const int32_t* offsets,
    const int32_t* totals, int previous_count, const float* params, const uint32_t* observation_counts, int count,
    float min_opacity, int32_t* labels) {
    using namespace gsn;
    if (!context || previous_count < 0 || count < 0 || !totals || !(min_opacity > 0.0f && min_opacity < 1.0f)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_outcomes: invalid arguments");
    }
    if (previous_count == 0) {
        return GSN_OK;
    }
    if (!decisions || !offsets || !labels || (count > 0 && (!params || !observation_counts))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_outcomes: null buffer");
    }
    Int4 t{totals[0], totals[1], totals[2], totals[3]};
    float logit = logf(min_opacity / (1.0f - min_opacity));
    outcome_labels_kernel<<<grid_for(previous_count, 256), 256, 0, context->stream>>>(
        reinterpret_cast<const Int4*>(decisions), reinterpret_cast<const Int4*>(offsets), t, previous_count, params,
        observation_counts, count, logit, labels);
    return check_launch("outcome_labels_kernel");
}
