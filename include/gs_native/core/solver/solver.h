// Solver (reference src/core/solver/solver.h:11-68): fused L1+SSIM image gradient, splat
// backward, projection statistics, chain rule and dense Adam, over the C ABI.
#pragma once

#include "core/types/buffers.h"

#include <cstdint>

namespace gs_native {

class Renderer;

class Solver {
public:
    Solver() = default;
    ~Solver();

    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    bool Iterate(
        const CameraParams& camera,
        const Renderer& renderer,
        const float* image,
        const uint8_t* target,
        GaussianBuffers& gaussians,
        float position_learning_rate,
        int sh_degree);

    // Addition: Iterate fused with the projection of the updated parameters for next_camera into
    // next_renderer's buffers (gsn_gaussian_gradients_adam_project); next_renderer then renders
    // that view with RenderProjected, provided nothing changes the parameters or the SH degree in
    // between. next_renderer must not be renderer.
    bool IterateAndProject(
        const CameraParams& camera,
        const Renderer& renderer,
        const float* image,
        const uint8_t* target,
        GaussianBuffers& gaussians,
        float position_learning_rate,
        int sh_degree,
        const CameraParams& next_camera,
        Renderer& next_renderer);

    bool ResetOpacityHistory();

    bool RemapGaussians(
        const int* parent_ids,
        int gaussian_count);

    const float* GetProjectionStatistics() const {
        return projection_statistics_;
    }

    const float* GetAbsoluteProjectionStatistics() const {
        return absolute_projection_statistics_;
    }

private:
    bool Step(
        const CameraParams& camera,
        const Renderer& renderer,
        const float* image,
        const uint8_t* target,
        GaussianBuffers& gaussians,
        float position_learning_rate,
        int sh_degree,
        const CameraParams* next_camera,
        Renderer* next_renderer);
    bool ComputeImageGradients(const CameraParams& camera, const float* image, const uint8_t* target);

    float* image_gradients_ = nullptr;
    int64_t pixel_capacity_ = 0;

    float* projection_statistics_ = nullptr;
    float* absolute_projection_statistics_ = nullptr;
    float* moments_ = nullptr;  // Adam moments, blocked (gsn_moment_floats floats; see gsn_hip.h)
    int gaussian_capacity_ = 0;

    uint64_t iteration_ = 0;
};

} // namespace gs_native
