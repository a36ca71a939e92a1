// DensityController over the C ABI (reference src/core/density/density_controller.cu:312-647).
#include "core/density/density_controller.h"

#include "device_memory.h"
#include "util/device.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <utility>
#include <vector>

namespace gs_native {

DensityController::~DensityController() {
    DeviceFree(spare_gaussian_parameters_);
    DeviceFree(history_parent_ids_);
    DeviceFree(parent_ids_);
    DeviceFree(child_flags_);
    DeviceFree(gradient_sums_);
    DeviceFree(absolute_gradient_sums_);
    DeviceFree(observation_counts_);
    DeviceFree(density_counts_);
    DeviceFree(density_offsets_);
    DeviceFree(feature_state_);
    DeviceFree(features_);
    DeviceFree(logits_);
    DeviceFree(control_);
}

// Grows the accumulators (values kept, new entries zero) and the decision buffers.
bool DensityController::Reserve(int capacity) {
    if (capacity <= capacity_) {
        return true;
    }
    gsn_context* context = DefaultContext();
    float* sums = nullptr;
    float* absolute = nullptr;
    uint32_t* counts = nullptr;
    int* decisions = nullptr;
    int* offsets = nullptr;
    size_t n = static_cast<size_t>(capacity);
    bool ok = context && DeviceAllocateCount(&sums, n) && DeviceAllocateCount(&absolute, n) &&
        DeviceAllocateCount(&counts, n) && DeviceAllocateCount(&decisions, 4 * (n + 1)) &&
        DeviceAllocateCount(&offsets, 4 * (n + 1));
    if (ok) {
        ok = Check(gsn_density_clear(context, sums, absolute, counts, capacity), "gsn_density_clear") &&
            Check(gsn_context_synchronize(context), "gsn_context_synchronize");
    }
    if (ok && capacity_ > 0) {
        // Decisions and offsets are copied too: ControlDensity grows the buffers between the
        // count pass that fills them and the write pass that reads them.
        size_t old = static_cast<size_t>(capacity_);
        ok = hipMemcpy(sums, gradient_sums_, old * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(absolute, absolute_gradient_sums_, old * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(counts, observation_counts_, old * sizeof(uint32_t), hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(decisions, density_counts_, 4 * (old + 1) * sizeof(int), hipMemcpyDeviceToDevice) == hipSuccess &&
            hipMemcpy(offsets, density_offsets_, 4 * (old + 1) * sizeof(int), hipMemcpyDeviceToDevice) == hipSuccess;
    }
    if (!ok) {
        DeviceFree(sums);
        DeviceFree(absolute);
        DeviceFree(counts);
        DeviceFree(decisions);
        DeviceFree(offsets);
        return false;
    }
    DeviceFree(gradient_sums_);
    DeviceFree(absolute_gradient_sums_);
    DeviceFree(observation_counts_);
    DeviceFree(density_counts_);
    DeviceFree(density_offsets_);
    gradient_sums_ = sums;
    absolute_gradient_sums_ = absolute;
    observation_counts_ = counts;
    density_counts_ = decisions;
    density_offsets_ = offsets;
    capacity_ = capacity;
    return true;
}

// control_ mirrors the reference's control block: [0] count, [1] next count, [2] status,
// [3] splits, [4] survivors, [5] clones, [6] kept splits, [7] previous count.
bool DensityController::WriteControl(int count, const int* totals, int previous_count) {
    int block[8] = {count, count, 0, 0, 0, 0, 0, previous_count};
    if (totals) {
        block[3] = totals[2];
        block[4] = totals[0];
        block[5] = totals[1];
        block[6] = totals[3];
    }
    return gsn_memcpy_to_device(control_, block, sizeof(block)) == GSN_OK;
}

bool DensityController::Initialize(int gaussian_count, int capacity) {
    if (gaussian_count < 0 || capacity < gaussian_count || capacity < 0 || capacity > INT_MAX / 6) {
        return false;
    }
    if (!DefaultContext()) {
        return false;
    }
    if (!control_ && !DeviceAllocateCount(&control_, 16)) {
        return false;
    }
    if (!Reserve(std::max(capacity, 1))) {
        return false;
    }
    if (!Check(gsn_density_clear(DefaultContext(), gradient_sums_, absolute_gradient_sums_, observation_counts_, capacity_),
            "gsn_density_clear")) {
        return false;
    }
    parameter_stride_ = capacity;
    seed_ = 0;
    counter_ = 0;
    return WriteControl(gaussian_count, nullptr, gaussian_count);
}

bool DensityController::AccumulateStatistics(const float* gradients, const float* radii, const int* gaussian_ids,
    int observation_count) {
    if (observation_count == 0) {
        return true;
    }
    if (!control_ || !gradients || !radii || observation_count < 0) {
        return false;
    }
    return Check(gsn_density_accumulate_observations(DefaultContext(), gradients, radii, gaussian_ids, observation_count,
                     parameter_stride_, gradient_sums_, observation_counts_),
        "gsn_density_accumulate_observations");
}

bool DensityController::AccumulateStatistics(const float* statistics, const float* absolute_statistics,
    int gaussian_count) {
    if (gaussian_count == 0) {
        return true;
    }
    if (!control_ || !statistics || !absolute_statistics || gaussian_count < 0 || gaussian_count > capacity_) {
        return false;
    }
    return Check(gsn_density_accumulate(DefaultContext(), statistics, absolute_statistics,
                     std::min(gaussian_count, parameter_stride_), gradient_sums_, absolute_gradient_sums_,
                     observation_counts_),
        "gsn_density_accumulate");
}

// Feature buffers sized for `count` Gaussians (stride feature_capacity_). They only grow right
// after a density step, when the state has just been cleared, so nothing is carried over.
bool DensityController::ReserveFeatures(int count) {
    if (count <= feature_capacity_) {
        return true;
    }
    int capacity = std::max(count + count / 4, 1);
    float* state = nullptr;
    float* features = nullptr;
    float* logits = nullptr;
    size_t n = static_cast<size_t>(capacity);
    if (!DeviceAllocateCount(&state, 6 * n) || !DeviceAllocateCount(&features, 16 * n) ||
        !DeviceAllocateCount(&logits, 3 * n) ||
        !Check(gsn_density_features_clear(DefaultContext(), state, capacity), "gsn_density_features_clear")) {
        DeviceFree(state);
        DeviceFree(features);
        DeviceFree(logits);
        return false;
    }
    DeviceFree(feature_state_);
    DeviceFree(features_);
    DeviceFree(logits_);
    feature_state_ = state;
    features_ = features;
    logits_ = logits;
    feature_capacity_ = capacity;
    return true;
}

bool DensityController::AccumulateFeatures(const float* statistics, const float* absolute_statistics,
    const float* depths, const float* conic_opacities, int gaussian_count) {
    if (!proposal_ || gaussian_count == 0) {
        return true;
    }
    if (!control_ || gaussian_count < 0 || !ReserveFeatures(gaussian_count)) {
        return false;
    }
    return Check(gsn_density_features_accumulate(DefaultContext(), statistics, absolute_statistics, depths,
                     conic_opacities, gaussian_count, feature_state_, feature_capacity_),
        "gsn_density_features_accumulate");
}

bool DensityController::ResetOpacity(float* parameters, int stride, float max_opacity) {
    if (!control_ || stride != parameter_stride_ || (!parameters && stride > 0) || !std::isfinite(max_opacity) ||
        max_opacity <= 0.0f || max_opacity >= 1.0f) {
        return false;
    }
    float max_opacity_parameter = std::log(max_opacity / (1.0f - max_opacity));
    return Check(gsn_density_reset_opacity(DefaultContext(), parameters, stride, stride, max_opacity_parameter),
        "gsn_density_reset_opacity");
}

bool DensityController::ControlDensity(GaussianBuffers& gaussians, float scene_extent, float gradient_threshold,
    float percent_dense, float min_opacity, float max_screen_size, unsigned long long seed, int* changes,
    float absolute_gradient_threshold) {
    int count = gaussians.gaussian_count_;
    float* parameters = gaussians.gaussians_;
    gsn_context* context = DefaultContext();
    if (!context || !control_ || (!parameters && count > 0) || (parameters && parameters == spare_gaussian_parameters_) ||
        count < 0 || count != parameter_stride_ || !std::isfinite(scene_extent) || scene_extent < 0 ||
        !std::isfinite(gradient_threshold) || gradient_threshold < 0 || !std::isfinite(percent_dense) ||
        percent_dense < 0 || !std::isfinite(min_opacity) || min_opacity < 0 || min_opacity > 1 ||
        !std::isfinite(max_screen_size) || max_screen_size < 0 || !std::isfinite(absolute_gradient_threshold) ||
        absolute_gradient_threshold < 0) {
        return false;
    }
    if (!Reserve(std::max(count, 1))) {
        return false;
    }
    int totals[4] = {0, 0, 0, 0};
    float split_threshold = percent_dense * scene_extent;
    bool proposed = proposal_ && proposal_->IsLoaded() && count > 0 && gradient_threshold > 0.0f &&
        split_threshold > 0.0f && scene_extent > 0.0f;
    if (proposed) {
        // Learned proposal: features -> logits -> decisions; pruning stays the rule's.
        if (!ReserveFeatures(count) ||
            !Check(gsn_density_features(context, parameters, count, gradient_sums_, absolute_gradient_sums_,
                       observation_counts_, feature_state_, feature_capacity_, gradient_threshold, split_threshold,
                       scene_extent, features_),
                "gsn_density_features") ||
            !proposal_->Predict(features_, count, logits_)) {
            return false;
        }
    }
    auto rule_counts = [&](int* out) {
        return Check(gsn_density_count(context, parameters, count, gradient_sums_, absolute_gradient_sums_,
                         observation_counts_, gradient_threshold, split_threshold, min_opacity, 0.1f * scene_extent,
                         max_screen_size > 0, absolute_gradient_threshold, density_counts_, density_offsets_, out),
            "gsn_density_count");
    };
    if (proposed) {
        // Growth guard: the rule's count first, then the proposal; a proposal adding more than
        // twice the rule's Gaussians (and more than 1% of the scene) is not followed.
        int rule_totals[4] = {0, 0, 0, 0};
        if (!rule_counts(rule_totals) ||
            !Check(gsn_density_count_proposed(context, parameters, count, gradient_sums_, absolute_gradient_sums_,
                       observation_counts_, logits_, split_threshold, min_opacity, 0.1f * scene_extent,
                       max_screen_size > 0, density_counts_, density_offsets_, totals),
                "gsn_density_count_proposed")) {
            return false;
        }
        int rule_added = rule_totals[0] + rule_totals[1] + 2 * rule_totals[3] - count;
        int proposed_added = totals[0] + totals[1] + 2 * totals[3] - count;
        if (proposed_added > std::max(2 * std::max(rule_added, 0), count / 100)) {
            ++proposal_fallbacks_;
            if (!rule_counts(totals)) {
                return false;
            }
        }
    }
    else if (!rule_counts(totals)) {
        return false;
    }
    int required = totals[0] + totals[1] + 2 * totals[3];
    if (changes) {
        changes[0] = totals[1] + 2 * totals[3];
        changes[1] = count - totals[0];
    }
    if (!Reserve(std::max(count, std::max(required, spare_gaussian_stride_)))) {
        return false;
    }
    if (required == 0) {
        DeviceFree(spare_gaussian_parameters_);
        spare_gaussian_parameters_ = nullptr;
        spare_gaussian_stride_ = 0;
    }
    else if (!spare_gaussian_parameters_ || required > spare_gaussian_stride_) {
        float* replacement = nullptr;
        if (!DeviceAllocateCount(&replacement, static_cast<size_t>(kMaxParamCount) * required)) {
            return false;
        }
        DeviceFree(spare_gaussian_parameters_);
        spare_gaussian_parameters_ = replacement;
        spare_gaussian_stride_ = required;
    }
    if (required > 0) {
        int* history = nullptr;
        int* parents = nullptr;
        float* flags = nullptr;
        if (!DeviceAllocateCount(&history, required) || !DeviceAllocateCount(&parents, required) ||
            !DeviceAllocateCount(&flags, required)) {
            DeviceFree(history);
            DeviceFree(parents);
            DeviceFree(flags);
            return false;
        }
        DeviceFree(history_parent_ids_);
        DeviceFree(parent_ids_);
        DeviceFree(child_flags_);
        history_parent_ids_ = history;
        parent_ids_ = parents;
        child_flags_ = flags;
    }
    if (seed != seed_) {
        seed_ = seed;
        counter_ = 0;
    }
    bool ok = Check(gsn_density_write(context, parameters, count, density_counts_, density_offsets_, totals, seed_,
                        counter_, spare_gaussian_parameters_, required, history_parent_ids_, parent_ids_,
                        child_flags_),
                  "gsn_density_write") &&
        Check(gsn_density_clear(context, gradient_sums_, absolute_gradient_sums_, observation_counts_, count),
            "gsn_density_clear") &&
        (!feature_state_ ||
            Check(gsn_density_features_clear(context, feature_state_, feature_capacity_), "gsn_density_features_clear"));
    if (!ok) {
        return false;
    }
    ++counter_;
    parameter_stride_ = required;
    std::swap(gaussians.gaussians_, spare_gaussian_parameters_);
    gaussians.gaussian_count_ = required;
    spare_gaussian_stride_ = count;
    return WriteControl(required, totals, count);
}

} // namespace gs_native
