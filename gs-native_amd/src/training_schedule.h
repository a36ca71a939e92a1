// The training schedule of gs_native's GaussianSplatting::Train as a table of pure functions of
// the iteration number (reference src/core/gaussian_splatting.cpp:136-185 fixes the numbers).
// Keeping the schedule apart from the loop lets the C++ driver and the Python trainer read the
// same decisions and lets a test enumerate them without a GPU.
#pragma once

#include <algorithm>
#include <cmath>

namespace gs_native {
namespace schedule {

constexpr int kShDegreeStep = 1000;      // one more SH band every 1000 iterations ...
constexpr int kMaxShDegree = 3;          // ... up to degree 3
constexpr int kDecayIterations = 30000;  // position learning rate: log-linear over 30000
constexpr float kInitialPositionRate = 0.00016f;
constexpr float kFinalPositionRate = 0.0000016f;
constexpr int kStatisticsUntil = 15000;  // statistics and density control stop here
constexpr int kDensityAfter = 500;       // first density step after this iteration
constexpr int kDensityInterval = 100;
constexpr int kOpacityInterval = 500;    // opacity cap ...
constexpr int kOpacityResetInterval = 3000;  // ... which is a reset to 0.01 at these
constexpr float kOpacityCap = 0.8f;
constexpr float kOpacityReset = 0.01f;
constexpr int kScreenPruneAfter = 3000;  // size pruning from here on
constexpr float kMaxScreenSize = 20.0f;
// Density thresholds passed to DensityController::ControlDensity.
constexpr float kGradientThreshold = 0.0002f;
constexpr float kPercentDense = 0.01f;
constexpr float kMinOpacity = 0.005f;
constexpr float kAbsoluteGradientThreshold = 0.0008f;

// What iteration `iteration` (1-based) renders with.
struct Step {
    int sh_degree;
    float position_learning_rate;  // already scaled by the scene extent
};

inline Step StepFor(int iteration, float scene_extent) {
    float progress = std::min(1.0f, static_cast<float>(iteration) / static_cast<float>(kDecayIterations));
    float log_rate = std::log(kInitialPositionRate) * (1.0f - progress) + std::log(kFinalPositionRate) * progress;
    return {std::min(kMaxShDegree, iteration / kShDegreeStep), scene_extent * std::exp(log_rate)};
}

// What happens after `completed` iterations have finished.
struct Upkeep {
    bool accumulate = false;       // add this view's statistics to the density accumulators
    bool control_density = false;  // clone / split / prune
    float max_screen_size = 0.0f;  // > 0 enables pruning of oversized Gaussians
    float opacity_cap = 0.0f;      // > 0: cap every opacity at this value and clear its moments
};

inline Upkeep UpkeepAfter(int completed) {
    Upkeep upkeep;
    if (completed >= kStatisticsUntil) {
        return upkeep;
    }
    upkeep.accumulate = true;
    upkeep.control_density = completed > kDensityAfter && completed % kDensityInterval == 0;
    upkeep.max_screen_size = completed > kScreenPruneAfter ? kMaxScreenSize : 0.0f;
    if (completed % kOpacityInterval == 0) {
        upkeep.opacity_cap = completed % kOpacityResetInterval == 0 ? kOpacityReset : kOpacityCap;
    }
    return upkeep;
}

} // namespace schedule
} // namespace gs_native
