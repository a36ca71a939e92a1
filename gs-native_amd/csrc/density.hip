// Density control: clone / split / prune decisions, int4 stream compaction and the compacted
// parameter write, split jitter, statistics accumulation and opacity reset.
// Follows reference src/core/density/density_controller.cu:39-308.
//
// Every float operation that decides a mask or produces an output parameter runs with FP
// contraction off and the shared FP32 transcendental functions, so decisions and the written
// parameters equal the host restatement bit for bit.
//
// Split jitter: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11) keyed by the 64-bit seed, counter (i / 4, call counter), then Box-Muller
// on each 32-bit pair. Normal i of a call is component i % 4 of block i / 4, so the stream does
// not depend on the launch shape. (The reference draws from cuRAND's generator instead; the
// jitter values therefore differ, the distribution does not.)
#include "gsn_internal.h"

namespace gsn {

GSN_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t m0 = 0xD2511F53u, m1 = 0xCD9E8D57u;
    const uint32_t w0 = 0x9E3779B9u, w1 = 0xBB67AE85u;
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = static_cast<uint64_t>(m0) * c[0];
        uint64_t p1 = static_cast<uint64_t>(m1) * c[2];
        uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += w0;
        k1 += w1;
    }
}

// Uniform in (0, 1): 24 high bits plus half an ulp, never 0 or 1.
GSN_HD float uniform_open(uint32_t x) {
#pragma clang fp contract(off)
    return static_cast<float>(x >> 8) * 0x1p-24f + 0x1p-25f;
}

GSN_HD float jitter_normal(uint64_t seed, uint64_t counter, int64_t index) {
#pragma clang fp contract(off)
    uint64_t block = static_cast<uint64_t>(index) >> 2;
    uint32_t c[4] = {static_cast<uint32_t>(block), static_cast<uint32_t>(block >> 32), static_cast<uint32_t>(counter),
        static_cast<uint32_t>(counter >> 32)};
    philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
    int component = static_cast<int>(index & 3);
    int pair = component >> 1;
    float u1 = uniform_open(c[2 * pair]);
    float u2 = uniform_open(c[2 * pair + 1]);
    float radius = sqrtf(-2.0f * det_logf(u1));
    float s, co;
    det_sincosf(6.28318530717958647692f * u2, &s, &co);
    return (component & 1) ? radius * s : radius * co;
}

namespace {

__global__ __launch_bounds__(256) void density_count_kernel(const float* __restrict__ params,
    const float* __restrict__ gradient_sums, const float* __restrict__ absolute_gradient_sums,
    const uint32_t* __restrict__ observation_counts, Int4* decisions, int count, float gradient_threshold,
    float split_threshold, float min_opacity, float max_world_size, int prune_large,
    float absolute_gradient_threshold, const float* __restrict__ proposal) {
#pragma clang fp contract(off)
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id > count) {
        return;
    }
    if (id == count) {
        decisions[id] = Int4{0, 0, 0, 0};
        return;
    }
    const int64_t n = count;
    float opacity = 1.0f / (1.0f + det_expf(-params[kOpacityIndex * n + id]));
    uint32_t observations = observation_counts[id];
    float gradient = observations ? gradient_sums[id] / static_cast<float>(observations) : 0.0f;
    if (gradient != gradient) {
        gradient = 0.0f;
    }
    float scale[3];
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        scale[axis] = det_expf(params[(kLogScaleIndex + axis) * n + id]);
    }
    float max_scale = fmaxf(scale[0], fmaxf(scale[1], scale[2]));
    if (scale[0] != scale[0] || scale[1] != scale[1] || scale[2] != scale[2]) {
        max_scale = bits_to_float(0x7fc00000u);
    }
    bool clone, split;
    if (proposal) {
        // Learned proposal: the class with the largest logit (ties and NaN favour keep, then
        // clone) replaces the gradient-threshold selection; pruning below stays rule-based.
        float keep_logit = proposal[3 * static_cast<int64_t>(id)];
        float clone_logit = proposal[3 * static_cast<int64_t>(id) + 1];
        float split_logit = proposal[3 * static_cast<int64_t>(id) + 2];
        clone = clone_logit > keep_logit && clone_logit >= split_logit;
        split = split_logit > keep_logit && split_logit > clone_logit;
    }
    else {
        bool selected = gradient >= gradient_threshold;
        clone = selected && max_scale <= split_threshold;
        float absolute_gradient =
            observations ? absolute_gradient_sums[id] / static_cast<float>(observations) : 0.0f;
        bool split_selected =
            absolute_gradient_threshold > 0.0f ? absolute_gradient >= absolute_gradient_threshold : selected;
        split = split_selected && max_scale > split_threshold;
    }
    float remaining = max_scale;
    if (split && prune_large) {
        float x = det_expf(det_logf(scale[0] / 1.6f));
        float y = det_expf(det_logf(scale[1] / 1.6f));
        float z = det_expf(det_logf(scale[2] / 1.6f));
        remaining = fmaxf(x, fmaxf(y, z));
        if (x != x || y != y || z != z) {
            remaining = bits_to_float(0x7fc00000u);
        }
    }
    bool remove = opacity < min_opacity || (prune_large && remaining > max_world_size);
    decisions[id] = Int4{!split && !remove, clone && !remove, split, split && !remove};
}

__global__ __launch_bounds__(256) void density_write_kernel(const float* __restrict__ params,
    const Int4* __restrict__ decisions, const Int4* __restrict__ offsets, Int4 totals, int count, uint64_t seed,
    uint64_t counter, float* __restrict__ out, int out_count, int* history_parent_ids, int* parent_ids,
    float* child_flags) {
#pragma clang fp contract(off)
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    Int4 d = decisions[id];
    Int4 o = offsets[id];
    bool split = d.w != 0;
    if (!d.x && !d.y && !split) {
        return;
    }
    const int64_t n = count;
    const int64_t m = out_count;
    float scale[3] = {0.0f, 0.0f, 0.0f};
    float rotation[9] = {};
    if (split) {
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            scale[axis] = det_expf(params[(kLogScaleIndex + axis) * n + id]);
        }
        float x = params[(kRotationIndex + 0) * n + id];
        float y = params[(kRotationIndex + 1) * n + id];
        float z = params[(kRotationIndex + 2) * n + id];
        float angle = sqrtf(x * x + y * y + z * z);
        float a = 1.0f, b = 0.5f;
        if (angle > 1e-4f) {
            a = det_sinf(angle) / angle;
            float half_sinc = det_sinf(angle * 0.5f) / (angle * 0.5f);
            b = half_sinc * half_sinc * 0.5f;
        }
        else {
            float square = angle * angle;
            a -= square / 6.0f;
            b -= square / 24.0f;
        }
        rotation[0] = 1.0f - b * (y * y + z * z);
        rotation[1] = b * x * y - a * z;
        rotation[2] = b * x * z + a * y;
        rotation[3] = b * x * y + a * z;
        rotation[4] = 1.0f - b * (x * x + z * z);
        rotation[5] = b * y * z - a * x;
        rotation[6] = b * x * z - a * y;
        rotation[7] = b * y * z + a * x;
        rotation[8] = 1.0f - b * (x * x + y * y);
    }
    int children = split ? 2 : d.x + d.y;
    for (int child = 0; child < children; ++child) {
        int destination = split ? totals.x + totals.y + child * totals.w + o.w
                                : (child == 0 && d.x ? o.x : totals.x + o.y);
        if (destination < 0 || destination >= out_count) {
            continue;  // inconsistent decisions/offsets from the caller: never write out of bounds
        }
        float normals[3] = {0.0f, 0.0f, 0.0f};
        if (split) {
#pragma unroll
            for (int axis = 0; axis < 3; ++axis) {
                int64_t sample = (static_cast<int64_t>(child) * totals.z + o.z) * 3 + axis;
                normals[axis] = jitter_normal(seed, counter, sample);
            }
        }
        for (int parameter = 0; parameter < kParamCount; ++parameter) {
            float value = params[parameter * n + id];
            if (split && parameter < kPositionIndex + 3) {
                float delta = 0.0f;
#pragma unroll
                for (int axis = 0; axis < 3; ++axis) {
                    float local = normals[axis] * scale[axis];
                    delta += rotation[parameter * 3 + axis] * local;
                }
                value += delta;
            }
            else if (split && parameter >= kLogScaleIndex && parameter < kLogScaleIndex + 3) {
                value = det_logf(scale[parameter - kLogScaleIndex] / 1.6f);
            }
            out[parameter * m + destination] = value;
        }
        if (history_parent_ids) {
            history_parent_ids[destination] = !split && child == 0 && d.x ? id : -1;
        }
        if (parent_ids) {
            parent_ids[destination] = id;
        }
        if (child_flags) {
            child_flags[destination] = (d.y || d.z) ? 1.0f : 0.0f;
        }
    }
}

__global__ __launch_bounds__(256) void density_normals_kernel(uint64_t seed, uint64_t counter, int64_t n, float* out) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        out[i] = jitter_normal(seed, counter, i);
    }
}

__global__ __launch_bounds__(256) void density_accumulate_kernel(const float* statistics,
    const float* absolute_statistics, int count, float* sums, float* absolute_sums, uint32_t* counts) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    sums[id] += statistics[id];
    absolute_sums[id] += absolute_statistics[id];
    counts[id] += static_cast<uint32_t>(statistics[static_cast<int64_t>(count) + id]);
}

__global__ __launch_bounds__(256) void density_observations_kernel(const float* gradients, const float* radii,
    const int* ids, int observations, int count, float* sums, uint32_t* counts) {
    int index = blockIdx.x * blockDim.x + threadIdx.x;
    if (index >= observations || !(radii[index] > 0.0f)) {
        return;
    }
    int id = ids ? ids[index] : index;
    if (id < 0 || id >= count) {
        return;
    }
    float x = gradients[2ll * index], y = gradients[2ll * index + 1];
    atomicAdd(sums + id, sqrtf(x * x + y * y));
    atomicAdd(counts + id, 1u);
}

__global__ __launch_bounds__(256) void density_reset_opacity_kernel(float* params, int stride, int count,
    float max_logit) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    int64_t index = static_cast<int64_t>(kOpacityIndex) * stride + id;
    params[index] = fminf(params[index], max_logit);
}

__global__ __launch_bounds__(256) void density_clear_kernel(float* sums, float* absolute_sums, uint32_t* counts,
    int count) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id < count) {
        sums[id] = 0.0f;
        absolute_sums[id] = 0.0f;
        counts[id] = 0u;
    }
}

} // namespace
} // namespace gsn

namespace gsn {
namespace {

int density_count(gsn_context* context, const float* params, int count, const float* gradient_sums,
    const float* absolute_gradient_sums, const uint32_t* observation_counts, float gradient_threshold,
    float split_threshold, float min_opacity, float max_world_size, int prune_large,
    float absolute_gradient_threshold, const float* proposal, int32_t* decisions, int32_t* offsets,
    int32_t* totals) {
    if (!context || count < 0 || !decisions || !offsets || !totals) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_count: invalid arguments");
    }
    if (count > 0 && (!params || !gradient_sums || !absolute_gradient_sums || !observation_counts)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_count: null buffer");
    }
    Int4* d = reinterpret_cast<Int4*>(decisions);
    Int4* o = reinterpret_cast<Int4*>(offsets);
    StageScope scope(context, kStageDensity);
    density_count_kernel<<<grid_for(static_cast<int64_t>(count) + 1, 256), 256, 0, context->stream>>>(params,
        gradient_sums, absolute_gradient_sums, observation_counts, d, count, gradient_threshold, split_threshold,
        min_opacity, max_world_size, prune_large, absolute_gradient_threshold, proposal);
    GSN_TRY(check_launch("density_count_kernel"));
    GSN_TRY(exclusive_scan_int4(context, d, o, static_cast<int64_t>(count) + 1));
    int* pinned = static_cast<int*>(context->pinned);
    GSN_HIP(hipMemcpyAsync(pinned, o + count, sizeof(Int4), hipMemcpyDeviceToHost, context->stream));
    GSN_HIP(hipStreamSynchronize(context->stream));
    for (int k = 0; k < 4; ++k) {
        totals[k] = pinned[k];
    }
    return GSN_OK;
}

} // namespace
} // namespace gsn

extern "C" int gsn_density_count(gsn_context* context, const float* params, int count, const float* gradient_sums,
    const float* absolute_gradient_sums, const uint32_t* observation_counts, float gradient_threshold,
    float split_threshold, float min_opacity, float max_world_size, int prune_large,
    float absolute_gradient_threshold, int32_t* decisions, int32_t* offsets, int32_t* totals) {
    return gsn::density_count(context, params, count, gradient_sums, absolute_gradient_sums, observation_counts,
        gradient_threshold, split_threshold, min_opacity, max_world_size, prune_large, absolute_gradient_threshold,
        nullptr, decisions, offsets, totals);
}

extern "C" int gsn_density_count_proposed(gsn_context* context, const float* params, int count,
    const float* gradient_sums, const float* absolute_gradient_sums, const uint32_t* observation_counts,
    const float* proposal_logits, float split_threshold, float min_opacity, float max_world_size, int prune_large,
    int32_t* decisions, int32_t* offsets, int32_t* totals) {
    if (count > 0 && !proposal_logits) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_count_proposed: null proposal logits");
    }
    return gsn::density_count(context, params, count, gradient_sums, absolute_gradient_sums, observation_counts,
        0.0f, split_threshold, min_opacity, max_world_size, prune_large, 0.0f, proposal_logits, decisions, offsets,
        totals);
}

extern "C" int gsn_density_write(gsn_context* context, const float* params, int count, const int32_t* decisions,
    const int32_t* offsets, const int32_t* totals, uint64_t seed, uint64_t counter, float* out_params, int out_count,
    int32_t* history_parent_ids, int32_t* parent_ids, float* child_flags) {
    using namespace gsn;
    if (!context || count < 0 || !totals || out_count < 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_write: invalid arguments");
    }
    Int4 t{totals[0], totals[1], totals[2], totals[3]};
    if (out_count != t.x + t.y + 2 * t.w) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_write: out_count %d != keep + clone + 2 * split (%d)",
            out_count, t.x + t.y + 2 * t.w);
    }
    if (count == 0 || out_count == 0) {
        return GSN_OK;
    }
    if (!params || !decisions || !offsets || !out_params) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_write: null buffer");
    }
    StageScope scope(context, kStageDensity);
    density_write_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(params,
        reinterpret_cast<const Int4*>(decisions), reinterpret_cast<const Int4*>(offsets), t, count, seed, counter,
        out_params, out_count, history_parent_ids, parent_ids, child_flags);
    return check_launch("density_write_kernel");
}

extern "C" int gsn_density_normals(gsn_context* context, uint64_t seed, uint64_t counter, int64_t n, float* out) {
    using namespace gsn;
    if (!context || n < 0 || (n > 0 && !out)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_normals: invalid arguments");
    }
    if (n == 0) {
        return GSN_OK;
    }
    density_normals_kernel<<<grid_for(n, 256), 256, 0, context->stream>>>(seed, counter, n, out);
    return check_launch("density_normals_kernel");
}

extern "C" int gsn_density_accumulate(gsn_context* context, const float* statistics, const float* absolute_statistics,
    int count, float* gradient_sums, float* absolute_gradient_sums, uint32_t* observation_counts) {
    using namespace gsn;
    if (!context || count < 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_accumulate: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    if (!statistics || !absolute_statistics || !gradient_sums || !absolute_gradient_sums || !observation_counts) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_accumulate: null buffer");
    }
    StageScope scope(context, kStageAccumulate);
    density_accumulate_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(statistics, absolute_statistics, count,
        gradient_sums, absolute_gradient_sums, observation_counts);
    return check_launch("density_accumulate_kernel");
}

extern "C" int gsn_density_accumulate_observations(gsn_context* context, const float* gradients, const float* radii,
    const int32_t* gaussian_ids, int observations, int count, float* gradient_sums, uint32_t* observation_counts) {
    using namespace gsn;
    if (!context || observations < 0 || count < 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_accumulate_observations: invalid arguments");
    }
    if (observations == 0) {
        return GSN_OK;
    }
    if (!gradients || !radii || !gradient_sums || !observation_counts) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_accumulate_observations: null buffer");
    }
    density_observations_kernel<<<grid_for(observations, 256), 256, 0, context->stream>>>(gradients, radii,
        gaussian_ids, observations, count, gradient_sums, observation_counts);
    return check_launch("density_observations_kernel");
}

extern "C" int gsn_density_reset_opacity(gsn_context* context, float* params, int stride, int count,
    float max_opacity_logit) {
    using namespace gsn;
    if (!context || stride < 0 || count < 0 || count > stride || (count > 0 && !params)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_reset_opacity: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    StageScope scope(context, kStageDensity);
    density_reset_opacity_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(params, stride, count,
        max_opacity_logit);
    return check_launch("density_reset_opacity_kernel");
}

extern "C" int gsn_density_clear(gsn_context* context, float* gradient_sums, float* absolute_gradient_sums,
    uint32_t* observation_counts, int count) {
    using namespace gsn;
    if (!context || count < 0 || (count > 0 && (!gradient_sums || !absolute_gradient_sums || !observation_counts))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_density_clear: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    StageScope scope(context, kStageDensity);
    density_clear_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(gradient_sums, absolute_gradient_sums,
        observation_counts, count);
    return check_launch("density_clear_kernel");
}
