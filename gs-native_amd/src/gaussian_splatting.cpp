// GaussianSplatting over the gs_native classes (reference src/core/gaussian_splatting.cpp:
// 72-344). Differences: training images are uploaded once at Initialize and stay resident
// (the reference copies the target every iteration), and no per-iteration log line is printed
// unless GS_NATIVE_LOG=info.
#include "core/gaussian_splatting.h"

#include "core/knn/knn.h"
#include "device_memory.h"
#include "io/data_io.h"
#include "log.h"
#include "util/device.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <random>

namespace gs_native {

namespace {

Vector3f CameraCentre(const Camera& camera) {
    Vector3f centre = Vector3f::Zero();
    for (int axis = 0; axis < 3; ++axis) {
        for (int row = 0; row < 3; ++row) {
            centre[axis] -= camera.R_c_w_(row, axis) * camera.t_c_w_[row];
        }
    }
    return centre;
}

} // namespace

GaussianSplatting::~GaussianSplatting() {
    DeviceFree(gaussians_.gaussians_);
    DeviceFree(image_);
    for (uint8_t* target : targets_) {
        DeviceFree(target);
    }
}

bool GaussianSplatting::Initialize(const Scene& scene) {
    GS_NATIVE_INFO("Initializing scene: training cameras %zu, test cameras %zu, points %zu.",
        scene.train_cameras_.size(), scene.test_cameras_.size(), scene.point_cloud_.positions_.size());
    if (gaussians_.gaussians_ || scene.train_cameras_.empty() ||
        scene.train_cameras_.size() != scene.train_images_.size() || scene.point_cloud_.positions_.empty() ||
        scene.point_cloud_.positions_.size() != scene.point_cloud_.colors_.size() ||
        scene.point_cloud_.positions_.size() > static_cast<size_t>(std::numeric_limits<int>::max()) || !DefaultContext()) {
        return false;
    }
    size_t pixels = 0;
    for (size_t i = 0; i < scene.train_cameras_.size(); ++i) {
        const Camera& camera = scene.train_cameras_[i];
        const Image8& image = scene.train_images_[i];
        if (image.empty() || image.width != camera.image_size_.width || image.height != camera.image_size_.height ||
            image.data.size() != static_cast<size_t>(image.width) * image.height * 3) {
            return false;
        }
        pixels = std::max(pixels, static_cast<size_t>(image.width) * image.height);
    }
    scene_ = scene;
    Vector3f mean = Vector3f::Zero();
    for (const Camera& camera : scene_.train_cameras_) {
        Vector3f centre = CameraCentre(camera);
        for (int axis = 0; axis < 3; ++axis) {
            mean[axis] += centre[axis];
        }
    }
    for (int axis = 0; axis < 3; ++axis) {
        mean[axis] /= static_cast<float>(scene_.train_cameras_.size());
    }
    scene_extent_ = 0.0f;
    for (const Camera& camera : scene_.train_cameras_) {
        Vector3f centre = CameraCentre(camera);
        Vector3f d = MakeVector3f(centre[0] - mean[0], centre[1] - mean[1], centre[2] - mean[2]);
        scene_extent_ = std::max(scene_extent_, d.norm());
    }
    scene_extent_ *= 1.1f;
    if (scene_extent_ == 0.0f) {
        for (const Vector3f& point : scene_.point_cloud_.positions_) {
            Vector3f d = MakeVector3f(point[0] - mean[0], point[1] - mean[1], point[2] - mean[2]);
            scene_extent_ = std::max(scene_extent_, d.norm());
        }
        scene_extent_ = std::max(scene_extent_, 1e-6f);
    }
    DeviceFree(image_);
    image_ = nullptr;
    for (uint8_t* target : targets_) {
        DeviceFree(target);
    }
    targets_.clear();
    pixel_capacity_ = 0;
    if (!DeviceAllocateCount(&image_, pixels * 3)) {
        return false;
    }
    for (const Image8& image : scene_.train_images_) {
        uint8_t* target = nullptr;
        if (!DeviceAllocateCount(&target, image.data.size()) ||
            gsn_memcpy_to_device(target, image.data.data(), image.data.size()) != GSN_OK) {
            DeviceFree(target);
            return false;
        }
        targets_.push_back(target);
    }
    pixel_capacity_ = pixels;
    camera_order_.resize(scene_.train_cameras_.size());
    std::iota(camera_order_.begin(), camera_order_.end(), 0);
    camera_cursor_ = static_cast<int>(camera_order_.size());
    completed_ = 0;
    sh_degree_ = 0;
    if (!InitializeGaussians()) {
        GS_NATIVE_ERROR("Gaussian initialization failed.");
        return false;
    }
    return true;
}

bool GaussianSplatting::InitializeGaussians() {
    const PointCloud& cloud = scene_.point_cloud_;
    int count = static_cast<int>(cloud.positions_.size());
    std::vector<float> distances = ComputeMeanDist2Cuda(cloud.positions_);
    if (distances.size() != cloud.positions_.size()) {
        return false;
    }
    std::vector<float> parameters(static_cast<size_t>(count) * kMaxParamCount, 0.0f);
    for (int id = 0; id < count; ++id) {
        float scale = 0.5f * std::log(std::max(distances[id], 1e-7f));
        for (int axis = 0; axis < 3; ++axis) {
            parameters[static_cast<size_t>(axis) * count + id] = cloud.positions_[id][axis];
            parameters[static_cast<size_t>(kLogScaleIndex + axis) * count + id] = scale;
            parameters[static_cast<size_t>(kShIndex + axis * 16) * count + id] =
                (cloud.colors_[id][axis] - 0.5f) / 0.28209479177387814f;
        }
        parameters[static_cast<size_t>(kOpacityIndex) * count + id] = std::log(0.1f / 0.9f);
    }
    float* replacement = nullptr;
    if (!DeviceAllocateCount(&replacement, parameters.size())) {
        return false;
    }
    if (gsn_memcpy_to_device(replacement, parameters.data(), parameters.size() * sizeof(float)) != GSN_OK ||
        !density_.Initialize(count, count) || !solver_.RemapGaussians(nullptr, count)) {
        DeviceFree(replacement);
        return false;
    }
    DeviceFree(gaussians_.gaussians_);
    gaussians_.gaussians_ = replacement;
    gaussians_.gaussian_count_ = count;
    return true;
}

bool GaussianSplatting::Train(int iterations) {
    if (iterations < 0 || !gaussians_.gaussians_ || scene_.train_cameras_.empty()) {
        return false;
    }
    std::mt19937 generator(0);
    for (int iteration = 0; iteration < iterations; ++iteration) {
        if (gaussians_.gaussian_count_ == 0) {
            return false;
        }
        if (camera_cursor_ >= static_cast<int>(camera_order_.size())) {
            std::shuffle(camera_order_.begin(), camera_order_.end(), generator);
            camera_cursor_ = 0;
        }
        int index = camera_order_[camera_cursor_];
        CameraParams camera = scene_.train_cameras_[index].GetCameraParams();
        int next = completed_ + 1;
        sh_degree_ = std::min(3, next / 1000);
        float fraction = std::min(1.0f, static_cast<float>(next) / 30000.0f);
        float learning_rate =
            scene_extent_ * std::exp(std::log(0.00016f) * (1.0f - fraction) + std::log(0.0000016f) * fraction);
        if (!renderer_.Render(gaussians_, camera, background_, image_, sh_degree_)) {
            GS_NATIVE_ERROR("Iteration %d: rendering failed.", next);
            return false;
        }
        if (!solver_.Iterate(camera, renderer_, image_, targets_[index], gaussians_, learning_rate, sh_degree_)) {
            GS_NATIVE_ERROR("Iteration %d: solver failed.", next);
            return false;
        }
        completed_ = next;
        ++camera_cursor_;
        if (completed_ < 15000) {
            if (!density_.AccumulateStatistics(solver_.GetProjectionStatistics(),
                    solver_.GetAbsoluteProjectionStatistics(), gaussians_.gaussian_count_) ||
                !density_.AccumulateFeatures(solver_.GetProjectionStatistics(),
                    solver_.GetAbsoluteProjectionStatistics(), renderer_.GetDepths(),
                    reinterpret_cast<const float*>(renderer_.GetConicOpacities()), gaussians_.gaussian_count_)) {
                return false;
            }
            if (completed_ > 500 && completed_ % 100 == 0 && !ControlDensity()) {
                return false;
            }
            if (completed_ % 500 == 0) {
                float max_opacity = completed_ % 3000 == 0 ? 0.01f : 0.8f;
                if (!density_.ResetOpacity(gaussians_.gaussians_, gaussians_.gaussian_count_, max_opacity) ||
                    !solver_.ResetOpacityHistory()) {
                    return false;
                }
            }
        }
        GS_NATIVE_INFO("Iteration %d/%d, completed %d, SH %d, gaussians %d.", iteration + 1, iterations, completed_,
            sh_degree_, gaussians_.gaussian_count_);
    }
    return Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize");
}

bool GaussianSplatting::Render(const Camera& source, ImageF& image) {
    if (!gaussians_.gaussians_ || source.image_size_.width <= 0 || source.image_size_.height <= 0) {
        return false;
    }
    CameraParams camera = source.GetCameraParams();
    size_t pixels = static_cast<size_t>(camera.width_) * camera.height_;
    if (pixels > pixel_capacity_) {
        float* replacement = nullptr;
        if (!DeviceAllocateCount(&replacement, pixels * 3)) {
            return false;
        }
        DeviceFree(image_);
        image_ = replacement;
        pixel_capacity_ = pixels;
    }
    if (!renderer_.Render(gaussians_, camera, background_, image_, sh_degree_)) {
        return false;
    }
    image.width = camera.width_;
    image.height = camera.height_;
    image.data.resize(pixels * 3);
    return Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize") &&
        gsn_memcpy_to_host(image.data.data(), image_, pixels * 3 * sizeof(float)) == GSN_OK;
}

bool GaussianSplatting::Save(const std::string& path) {
    int count = gaussians_.gaussian_count_;
    if (count <= 0 || !gaussians_.gaussians_) {
        return false;
    }
    std::vector<float> parameters(static_cast<size_t>(count) * kMaxParamCount);
    if (!Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize") ||
        gsn_memcpy_to_host(parameters.data(), gaussians_.gaussians_, parameters.size() * sizeof(float)) != GSN_OK) {
        return false;
    }
    std::vector<Gaussian> gaussians(count);
    for (int id = 0; id < count; ++id) {
        Gaussian& g = gaussians[id];
        for (int axis = 0; axis < 3; ++axis) {
            g.position_[axis] = parameters[static_cast<size_t>(axis) * count + id];
            g.log_scale_[axis] = parameters[static_cast<size_t>(kLogScaleIndex + axis) * count + id];
            g.rotation_[axis] = parameters[static_cast<size_t>(kRotationIndex + axis) * count + id];
            for (int k = 0; k < 16; ++k) {
                g.sh_(k, axis) = parameters[static_cast<size_t>(kShIndex + axis * 16 + k) * count + id];
            }
        }
        g.opacity_logit_ = parameters[static_cast<size_t>(kOpacityIndex) * count + id];
    }
    return WriteGaussiansPly(path, gaussians);
}

bool GaussianSplatting::ControlDensity() {
    int changes[2] = {};
    if (!density_.ControlDensity(gaussians_, scene_extent_, 0.0002f, 0.01f, 0.005f, completed_ > 3000 ? 20.0f : 0.0f, 0,
            changes, 0.0008f)) {
        GS_NATIVE_ERROR("Iteration %d: density control failed.", completed_);
        return false;
    }
    if (!solver_.RemapGaussians(density_.GetHistoryParentIds(), gaussians_.gaussian_count_)) {
        GS_NATIVE_ERROR("Iteration %d: remapping Adam history after density control failed.", completed_);
        return false;
    }
    GS_NATIVE_INFO("Iteration %d: density control, gaussians -> %d, added %d, removed %d.", completed_,
        gaussians_.gaussian_count_, changes[0], changes[1]);
    return true;
}

} // namespace gs_native
