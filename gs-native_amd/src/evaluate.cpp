// Held-out evaluation (see include/gs_native/eval/evaluate.h; reference demo/render.cpp:107-155).
#include "eval/evaluate.h"

#include "core/render/renderer.h"
#include "device_memory.h"
#include "image_io.h"
#include "log.h"
#include "util/device.h"

#include <algorithm>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>

namespace gs_native {

bool UploadGaussians(const std::vector<Gaussian>& gaussians, GaussianBuffers& buffers) {
    const size_t n = gaussians.size();
    if (n == 0 || n > static_cast<size_t>(std::numeric_limits<int>::max())) {
        return false;
    }
    std::vector<float> rows(n * kMaxParamCount, 0.0f);
    auto row = [&](int index) { return rows.data() + static_cast<size_t>(index) * n; };
    for (size_t i = 0; i < n; ++i) {
        const Gaussian& g = gaussians[i];
        for (int axis = 0; axis < 3; ++axis) {
            row(kPositionIndex + axis)[i] = g.position_[axis];
            row(kLogScaleIndex + axis)[i] = g.log_scale_[axis];
            row(kRotationIndex + axis)[i] = g.rotation_[axis];
            for (int k = 0; k < 16; ++k) {
                row(kShIndex + 16 * axis + k)[i] = g.sh_(k, axis);
            }
        }
        row(kOpacityIndex)[i] = g.opacity_logit_;
    }
    float* device = nullptr;
    if (!DeviceAllocateCount(&device, rows.size()) ||
        gsn_memcpy_to_device(device, rows.data(), rows.size() * sizeof(float)) != GSN_OK) {
        DeviceFree(device);
        return false;
    }
    buffers.gaussians_ = device;
    buffers.gaussian_count_ = static_cast<int>(n);
    return true;
}

bool EvaluateViews(const GaussianBuffers& gaussians, const std::vector<Camera>& cameras,
    const std::vector<Image8>& targets, int sh_degree, std::vector<ViewScore>& scores, std::vector<Image8>* renders) {
    scores.clear();
    if (renders) {
        renders->clear();
    }
    gsn_context* context = DefaultContext();
    if (!context || !gaussians.gaussians_ || cameras.empty() || cameras.size() != targets.size()) {
        return false;
    }
    size_t largest = 0;
    for (const Image8& target : targets) {
        largest = std::max(largest, target.data.size());
    }
    float* image = nullptr;
    uint8_t* target_device = nullptr;
    uint8_t* quantized = nullptr;
    Renderer renderer;
    const float background[3] = {0.0f, 0.0f, 0.0f};
    bool ok = DeviceAllocateCount(&image, largest) && DeviceAllocateCount(&target_device, largest) &&
              DeviceAllocateCount(&quantized, largest);
    for (size_t v = 0; ok && v < cameras.size(); ++v) {
        const Camera& camera = cameras[v];
        const Image8& target = targets[v];
        if (target.width != camera.image_size_.width || target.height != camera.image_size_.height ||
            target.data.size() != static_cast<size_t>(target.width) * target.height * 3) {
            GS_NATIVE_ERROR("target of view %zu does not match its camera", v);
            ok = false;
            break;
        }
        double result[2] = {0.0, 0.0};
        ok = renderer.Render(gaussians, camera.GetCameraParams(), background, image, sh_degree) &&
             gsn_memcpy_to_device(target_device, target.data.data(), target.data.size()) == GSN_OK &&
             Check(gsn_image_quantized_error(context, target.width, target.height, image, target_device,
                       renders ? quantized : nullptr, result),
                 "gsn_image_quantized_error");
        if (!ok) {
            GS_NATIVE_ERROR("evaluating view %zu (%s) failed", v, camera.image_name_.c_str());
            break;
        }
        scores.push_back({camera.image_name_, result[1], result[0]});
        if (renders) {
            Image8 render;
            render.width = target.width;
            render.height = target.height;
            render.data.resize(target.data.size());
            ok = gsn_memcpy_to_host(render.data.data(), quantized, render.data.size()) == GSN_OK;
            renders->push_back(std::move(render));
        }
    }
    DeviceFree(image);
    DeviceFree(target_device);
    DeviceFree(quantized);
    return ok;
}

double MeanPsnr(const std::vector<ViewScore>& scores) {
    double sum = 0.0;
    for (const ViewScore& score : scores) {
        sum += score.psnr_db;
    }
    return scores.empty() ? 0.0 : sum / static_cast<double>(scores.size());
}

bool WriteEvaluation(const std::string& output_dir, const std::string& split, const std::vector<ViewScore>& scores,
    const std::vector<Image8>& renders, const std::vector<Image8>& targets) {
    namespace fs = std::filesystem;
    const fs::path destination = fs::path(output_dir) / split / "gs-native";
    std::error_code error;
    if (fs::exists(destination, error) && !fs::is_empty(destination, error)) {
        GS_NATIVE_ERROR("the output directory %s is not empty", destination.string().c_str());
        return false;
    }
    for (const char* directory : {"renders", "gt", "comparison"}) {
        fs::create_directories(destination / directory, error);
        if (error) {
            return false;
        }
    }
    std::ofstream report(destination / "psnr.csv");
    report << "image,psnr_db\n" << std::setprecision(10);
    for (size_t v = 0; v < scores.size(); ++v) {
        char name[32];
        std::snprintf(name, sizeof(name), "%05zu.png", v);
        report << name << ',' << scores[v].psnr_db << '\n';
        if (v < renders.size() && v < targets.size()) {
            const Image8& render = renders[v];
            const Image8& target = targets[v];
            Image8 side;  // render | target, side by side
            side.width = render.width + target.width;
            side.height = render.height;
            side.data.resize(static_cast<size_t>(side.width) * side.height * 3);
            for (int y = 0; y < side.height; ++y) {
                std::copy_n(&render.data[static_cast<size_t>(y) * render.width * 3], render.width * 3,
                    &side.data[static_cast<size_t>(y) * side.width * 3]);
                std::copy_n(&target.data[static_cast<size_t>(y) * target.width * 3], target.width * 3,
                    &side.data[(static_cast<size_t>(y) * side.width + render.width) * 3]);
            }
            if (!WritePngRgb((destination / "renders" / name).string(), render) ||
                !WritePngRgb((destination / "gt" / name).string(), target) ||
                !WritePngRgb((destination / "comparison" / name).string(), side)) {
                GS_NATIVE_ERROR("writing the images of view %zu failed", v);
                return false;
            }
        }
    }
    report << "mean," << MeanPsnr(scores) << '\n';
    return static_cast<bool>(report);
}

} // namespace gs_native
