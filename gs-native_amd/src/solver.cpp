// Solver over the C ABI (reference src/core/solver/solver.cu:693-929). Gradients and Adam run
// as one fused per-Gaussian pass, so the [58][N] gradient buffer of the reference is not kept.
#include "core/solver/solver.h"

#include "core/render/renderer.h"
#include "device_memory.h"
#include "util/device.h"

#include <algorithm>
#include <cmath>

namespace gs_native {

Solver::~Solver() {
    DeviceFree(image_gradients_);
    DeviceFree(projection_statistics_);
    DeviceFree(absolute_projection_statistics_);
    DeviceFree(moments_);
}

bool Solver::Iterate(
    const CameraParams& camera,
    const Renderer& renderer,
    const float* image,
    const uint8_t* target,
    GaussianBuffers& gaussians,
    float position_learning_rate,
    int sh_degree) {
    return Step(camera, renderer, image, target, gaussians, position_learning_rate, sh_degree, nullptr, nullptr);
}

bool Solver::IterateAndProject(
    const CameraParams& camera,
    const Renderer& renderer,
    const float* image,
    const uint8_t* target,
    GaussianBuffers& gaussians,
    float position_learning_rate,
    int sh_degree,
    const CameraParams& next_camera,
    Renderer& next_renderer) {
    if (&next_renderer == &renderer) {
        return false;
    }
    return Step(camera, renderer, image, target, gaussians, position_learning_rate, sh_degree, &next_camera,
        &next_renderer);
}

bool Solver::Step(
    const CameraParams& camera,
    const Renderer& renderer,
    const float* image,
    const uint8_t* target,
    GaussianBuffers& gaussians,
    float position_learning_rate,
    int sh_degree,
    const CameraParams* next_camera,
    Renderer* next_renderer) {
    gsn_context* context = DefaultContext();
    if (!context || !ComputeImageGradients(camera, image, target)) {
        return false;
    }
    int count = gaussians.gaussian_count_;
    if (count == 0) {
        return true;
    }
    int64_t pixels = static_cast<int64_t>(camera.width_) * camera.height_;
    int64_t tiles = static_cast<int64_t>(camera.tile_count_x_) * camera.tile_count_y_;
    if (count < 0 || !gaussians.gaussians_ || !image || !image_gradients_ || pixels > pixel_capacity_ ||
        camera.tile_count_x_ <= 0 || camera.tile_count_y_ <= 0 || tiles > renderer.GetTileCapacity() ||
        count > renderer.GetGaussianCapacity() || !renderer.GetTileGaussianOffsets() || !renderer.GetMeans() ||
        !renderer.GetConicOpacities() || !renderer.GetColors()) {
        return false;
    }
    if (count != gaussian_capacity_) {
        if (iteration_ != 0 || !RemapGaussians(nullptr, count)) {
            return false;
        }
    }
    if (!std::isfinite(position_learning_rate) || position_learning_rate < 0) {
        return false;
    }
    uint64_t iteration = iteration_ + 1;
    float first_correction = static_cast<float>(1.0 / (1.0 - std::pow(0.9, static_cast<double>(iteration))));
    float second_correction = static_cast<float>(1.0 / (1.0 - std::pow(0.999, static_cast<double>(iteration))));
    if (next_renderer) {
        ProjectionTargets next;
        bool ok = next_renderer->ProjectionBuffers(count, &next) &&
                  Check(gsn_gaussian_gradients_adam_project(context, AsCamera(&camera), count,
                            std::clamp(sh_degree, 0, 3), gaussians.gaussians_, renderer.GetTileGaussianIds(),
                            renderer.GetTileGaussianOffsets(), renderer.GetRecordCount(), renderer.GetRecordOffsets(),
                            reinterpret_cast<const int*>(renderer.GetTileRects()),
                            reinterpret_cast<const float*>(renderer.GetMeans()),
                            reinterpret_cast<const float*>(renderer.GetConicOpacities()),
                            reinterpret_cast<const float*>(renderer.GetColors()), image, image_gradients_, moments_,
                            position_learning_rate, first_correction, second_correction, projection_statistics_,
                            absolute_projection_statistics_, AsCamera(next_camera), next.means, next.conic_opacities,
                            next.colors, next.depths, next.tile_rects, next.record_offsets, nullptr, nullptr,
                            nullptr),
                      "gsn_gaussian_gradients_adam_project");
        if (ok) {
            iteration_ = iteration;
        }
        return ok;
    }
    bool ok = Check(gsn_gaussian_gradients_adam(context, AsCamera(&camera), count, std::clamp(sh_degree, 0, 3),
                        gaussians.gaussians_, renderer.GetTileGaussianIds(), renderer.GetTileGaussianOffsets(),
                        renderer.GetRecordCount(), renderer.GetRecordOffsets(),
                        reinterpret_cast<const int*>(renderer.GetTileRects()),
                        reinterpret_cast<const float*>(renderer.GetMeans()),
                        reinterpret_cast<const float*>(renderer.GetConicOpacities()),
                        reinterpret_cast<const float*>(renderer.GetColors()), image, image_gradients_, moments_,
                        position_learning_rate, first_correction, second_correction, projection_statistics_,
                        absolute_projection_statistics_),
        "gsn_gaussian_gradients_adam");
    if (ok) {
        iteration_ = iteration;
    }
    return ok;
}

bool Solver::ComputeImageGradients(const CameraParams& camera, const float* image, const uint8_t* target) {
    gsn_context* context = DefaultContext();
    if (!context || !image || !target || camera.width_ <= 0 || camera.height_ <= 0) {
        return false;
    }
    int64_t pixels = static_cast<int64_t>(camera.width_) * camera.height_;
    if (pixels > pixel_capacity_) {
        DeviceFree(image_gradients_);
        image_gradients_ = nullptr;
        pixel_capacity_ = 0;
        if (!DeviceAllocateCount(&image_gradients_, static_cast<size_t>(pixels) * 3)) {
            return false;
        }
        pixel_capacity_ = pixels;
    }
    return Check(gsn_image_gradients(context, camera.width_, camera.height_, image, target, image_gradients_),
        "gsn_image_gradients");
}

bool Solver::ResetOpacityHistory() {
    if (gaussian_capacity_ == 0) {
        return true;
    }
    gsn_context* context = DefaultContext();
    return context && Check(gsn_zero_rows(context, moments_, gaussian_capacity_, kOpacityIndex, 1), "gsn_zero_rows");
}

bool Solver::RemapGaussians(const int* parent_ids, int gaussian_count) {
    gsn_context* context = DefaultContext();
    if (!context || gaussian_count < 0) {
        return false;
    }
    float* statistics = gaussian_count > 0 && gaussian_count <= gaussian_capacity_ ? projection_statistics_ : nullptr;
    float* absolute = gaussian_count > 0 && gaussian_count <= gaussian_capacity_ ? absolute_projection_statistics_ : nullptr;
    float* moments = nullptr;
    if (gaussian_count > 0) {
        size_t n = static_cast<size_t>(gaussian_count);
        bool ok = DeviceAllocateCount(&moments, gsn_moment_floats(gaussian_count));
        if (ok && !statistics) {
            ok = DeviceAllocateCount(&statistics, n * 2);
        }
        if (ok && !absolute) {
            ok = DeviceAllocateCount(&absolute, n * 2);
        }
        if (ok) {
            ok = Check(gsn_remap_history(context, parent_ids, gaussian_capacity_, gaussian_count, moments_, moments),
                     "gsn_remap_history") &&
                Check(gsn_context_synchronize(context), "gsn_context_synchronize");
        }
        if (!ok) {
            if (statistics != projection_statistics_) {
                DeviceFree(statistics);
            }
            if (absolute != absolute_projection_statistics_) {
                DeviceFree(absolute);
            }
            DeviceFree(moments);
            return false;
        }
    }
    if (statistics != projection_statistics_) {
        DeviceFree(projection_statistics_);
    }
    if (absolute != absolute_projection_statistics_) {
        DeviceFree(absolute_projection_statistics_);
    }
    DeviceFree(moments_);
    projection_statistics_ = statistics;
    absolute_projection_statistics_ = absolute;
    moments_ = moments;
    gaussian_capacity_ = gaussian_count;
    return true;
}

} // namespace gs_native
