// DensityController (reference src/core/density/density_controller.h:9-83): statistics
// accumulation, clone / split / prune with stream compaction, opacity reset, over the C ABI.
// ControlDensity swaps the caller's parameter buffer with the controller's spare, as the
// reference does. GetGaussianCount/GetStatus/GetParentIds return device pointers.
#pragma once

#include "core/density/proposal_network.h"
#include "core/types/buffers.h"

#include <cstddef>
#include <cstdint>

namespace gs_native {

class DensityController {
public:
    DensityController() = default;
    ~DensityController();
    DensityController(const DensityController&) = delete;
    DensityController& operator=(const DensityController&) = delete;

    bool Initialize(
        int gaussian_count,
        int capacity);

    bool AccumulateStatistics(
        const float* gradients,
        const float* radii,
        const int* gaussian_ids,
        int observation_count);

    bool AccumulateStatistics(
        const float* statistics,
        const float* absolute_statistics,
        int gaussian_count);

    bool ResetOpacity(
        float* parameters,
        int stride,
        float max_opacity = 0.01f);

    bool ControlDensity(
        GaussianBuffers& gaussians,
        float scene_extent,
        float gradient_threshold = 0.0002f,
        float percent_dense = 0.01f,
        float min_opacity = 0.005f,
        float max_screen_size = 0.0f,
        unsigned long long seed = 0,
        int* changes = nullptr,
        float absolute_gradient_threshold = 0.0f);

    const int* GetHistoryParentIds() const {
        return history_parent_ids_;
    }

    const int* GetGaussianCount() const {
        return control_;
    }

    const int* GetStatus() const {
        return control_ ? control_ + 2 : nullptr;
    }

    const int* GetParentIds() const {
        return parent_ids_;
    }

    const float* GetChildFlags() const {
        return child_flags_;
    }

    // Learned densification (additions, no counterpart in the reference). With a proposal set,
    // ControlDensity builds the 16 features of every Gaussian, runs the network and lets its
    // logits choose keep / clone / split (gsn_density_count_proposed); pruning, compaction,
    // jitter and the buffer swap are unchanged. nullptr restores the gradient-threshold rule.
    // The features need AccumulateFeatures after every AccumulateStatistics.
    void SetProposal(const ProposalNetwork* network) {
        proposal_ = network;
    }

    // Density steps in which the proposal was not followed because it added more than twice the
    // Gaussians the plain rule would (and more than 1% of the scene): the rule decided instead.
    int GetProposalFallbacks() const {
        return proposal_fallbacks_;
    }
    const ProposalNetwork* GetProposal() const {
        return proposal_;
    }

    bool AccumulateFeatures(
        const float* statistics,
        const float* absolute_statistics,
        const float* depths,
        const float* conic_opacities,
        int gaussian_count);

private:
    bool Reserve(int capacity);
    bool ReserveFeatures(int count);
    bool WriteControl(int count, const int* totals, int previous_count);

    float* spare_gaussian_parameters_ = nullptr;
    int* history_parent_ids_ = nullptr;
    int* parent_ids_ = nullptr;
    float* child_flags_ = nullptr;
    int spare_gaussian_stride_ = 0;

    float* gradient_sums_ = nullptr;
    float* absolute_gradient_sums_ = nullptr;
    uint32_t* observation_counts_ = nullptr;
    int* density_counts_ = nullptr;
    int* density_offsets_ = nullptr;
    int* control_ = nullptr;
    const ProposalNetwork* proposal_ = nullptr;
    int proposal_fallbacks_ = 0;
    float* feature_state_ = nullptr;  // [6][capacity_]
    float* features_ = nullptr;       // [capacity_][16]
    float* logits_ = nullptr;         // [capacity_][3]
    int feature_capacity_ = 0;
    int capacity_ = 0;
    int parameter_stride_ = 0;
    unsigned long long seed_ = 0;
    unsigned long long counter_ = 0;
};

} // namespace gs_native
