// GaussianSplatting (reference src/core/gaussian_splatting.h:14-53): initialisation from a point
// cloud (3-NN scales), the per-iteration training schedule, rendering and PLY export.
#pragma once

#include "core/density/density_controller.h"
#include "core/render/renderer.h"
#include "core/solver/solver.h"
#include "core/types/scene.h"

#include <cstddef>
#include <random>
#include <string>
#include <vector>

namespace gs_native {

class GaussianSplatting {
public:
    GaussianSplatting() = default;
    ~GaussianSplatting();

    GaussianSplatting(const GaussianSplatting&) = delete;
    GaussianSplatting& operator=(const GaussianSplatting&) = delete;

    bool Initialize(const Scene& scene);

    bool Train(int iterations);

    // Renders into image (float RGB, unclamped), like the reference's CV_32FC3 output.
    bool Render(const Camera& camera, ImageF& image);

    bool Save(const std::string& path);

    int GetGaussianCount() const {
        return gaussians_.gaussian_count_;
    }

    // Learned densification (addition): with a loaded network, density control takes keep /
    // clone / split from it (DensityController::SetProposal); nullptr restores the plain rule.
    void SetDensityProposal(const ProposalNetwork* network) {
        density_.SetProposal(network);
    }

private:
    // One training iteration on the next camera of the shuffled cycle, then the schedule's upkeep.
    bool Step();
    bool Upkeep();
    int NextCamera();
    bool EnsureImageCapacity(size_t pixels);
    void ReleaseImages();

    Scene scene_;
    GaussianBuffers gaussians_;

    // Two renderers: the gradient pass projects the updated parameters for the next camera into
    // the spare one, which then renders that view (Solver::IterateAndProject). next_view_ is the
    // (camera, count, SH degree) of that projection, or camera -1 when parameters changed since.
    Renderer renderers_[2];
    int current_ = 0;
    int next_view_ = -1;
    int next_count_ = 0;
    int next_degree_ = 0;
    Renderer& renderer() {
        return renderers_[current_];
    }
    Solver solver_;
    DensityController density_;

    float* image_ = nullptr;             // device HWC float render target
    size_t image_pixels_ = 0;
    std::vector<uint8_t*> targets_;      // device copies of the training images, uploaded once

    std::vector<int> cycle_;             // camera order of the current pass over the views
    size_t cycle_position_ = 0;
    std::mt19937 shuffler_;              // restarted from seed 0 by every Train call

    int completed_ = 0;
    int sh_degree_ = 0;
    float scene_extent_ = 0.0f;
    float background_[3] = {0.0f, 0.0f, 0.0f};
};

} // namespace gs_native
