// GaussianSplatting: the driver of the training path (API of reference
// src/core/gaussian_splatting.h:14-53). The schedule lives in training_schedule.h; this file
// holds the state that survives between iterations (device images, camera cycle, iteration
// count) and one function per iteration.
//
// Training images are uploaded once at Initialize and stay resident in device memory, so an
// iteration moves no image data over PCIe. The camera cycle reproduces the reference's order: a
// std::mt19937 seeded with 0 at the start of every Train call reshuffles the running permutation
// whenever a pass over the views ends.
#include "core/gaussian_splatting.h"

#include "core/knn/knn.h"
#include "device_memory.h"
#include "io/data_io.h"
#include "log.h"
#include "training_schedule.h"
#include "util/device.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace gs_native {

namespace {

constexpr float kShDc = 0.28209479177387814f;       // Y_0^0: colour = 0.5 + kShDc * dc
constexpr float kInitialOpacity = 0.1f;
constexpr float kMinMeanDistance2 = 1e-7f;

// Camera centre c = -R^T t of a world-to-camera pose.
Vector3f CentreOf(const Camera& camera) {
    Vector3f c = Vector3f::Zero();
    for (int row = 0; row < 3; ++row) {
        for (int axis = 0; axis < 3; ++axis) {
            c[axis] -= camera.R_c_w_(row, axis) * camera.t_c_w_[row];
        }
    }
    return c;
}

float Distance(const Vector3f& a, const Vector3f& b) {
    return MakeVector3f(a[0] - b[0], a[1] - b[1], a[2] - b[2]).norm();
}

// Scene extent: 1.1 x the largest distance of a training camera centre from their mean; when all
// centres coincide, the largest distance of a point from that mean (at least 1e-6).
float SceneExtent(const std::vector<Camera>& cameras, const std::vector<Vector3f>& points) {
    Vector3f mean = Vector3f::Zero();
    for (const Camera& camera : cameras) {
        Vector3f c = CentreOf(camera);
        for (int axis = 0; axis < 3; ++axis) {
            mean[axis] += c[axis];
        }
    }
    for (int axis = 0; axis < 3; ++axis) {
        mean[axis] /= static_cast<float>(cameras.size());
    }
    float spread = 0.0f;
    for (const Camera& camera : cameras) {
        spread = std::max(spread, Distance(CentreOf(camera), mean));
    }
    if (spread * 1.1f != 0.0f) {
        return spread * 1.1f;
    }
    float reach = 0.0f;
    for (const Vector3f& point : points) {
        reach = std::max(reach, Distance(point, mean));
    }
    return std::max(reach, 1e-6f);
}

// Initial [58][N] parameters: positions from the points, isotropic log-scales from the mean
// squared 3-NN distance (half its log), identity rotation, opacity 0.1, SH DC from the point
// colours and no higher bands.
std::vector<float> InitialParameters(const PointCloud& cloud, const std::vector<float>& mean_distance2) {
    const size_t n = cloud.positions_.size();
    std::vector<float> rows(n * kMaxParamCount, 0.0f);
    auto row = [&](int index) { return rows.data() + static_cast<size_t>(index) * n; };
    const float opacity_logit = std::log(kInitialOpacity / (1.0f - kInitialOpacity));
    for (size_t i = 0; i < n; ++i) {
        const float log_scale = 0.5f * std::log(std::max(mean_distance2[i], kMinMeanDistance2));
        for (int axis = 0; axis < 3; ++axis) {
            row(kPositionIndex + axis)[i] = cloud.positions_[i][axis];
            row(kLogScaleIndex + axis)[i] = log_scale;
            row(kShIndex + 16 * axis)[i] = (cloud.colors_[i][axis] - 0.5f) / kShDc;
        }
        row(kOpacityIndex)[i] = opacity_logit;
    }
    return rows;
}

// [58][N] parameters to per-Gaussian records (axis-angle kept; the PLY writer converts it).
std::vector<Gaussian> GaussiansFromRows(const std::vector<float>& rows, int count) {
    std::vector<Gaussian> out(static_cast<size_t>(count));
    auto at = [&](int index, int id) { return rows[static_cast<size_t>(index) * count + id]; };
    for (int id = 0; id < count; ++id) {
        Gaussian& g = out[static_cast<size_t>(id)];
        for (int axis = 0; axis < 3; ++axis) {
            g.position_[axis] = at(kPositionIndex + axis, id);
            g.log_scale_[axis] = at(kLogScaleIndex + axis, id);
            g.rotation_[axis] = at(kRotationIndex + axis, id);
            for (int k = 0; k < 16; ++k) {
                g.sh_(k, axis) = at(kShIndex + 16 * axis + k, id);
            }
        }
        g.opacity_logit_ = at(kOpacityIndex, id);
    }
    return out;
}

bool ImageMatches(const Image8& image, const Camera& camera) {
    return !image.empty() && image.width == camera.image_size_.width && image.height == camera.image_size_.height &&
           image.data.size() == static_cast<size_t>(image.width) * image.height * 3;
}

} // namespace

GaussianSplatting::~GaussianSplatting() {
    DeviceFree(gaussians_.gaussians_);
    ReleaseImages();
}

void GaussianSplatting::ReleaseImages() {
    DeviceFree(image_);
    image_ = nullptr;
    image_pixels_ = 0;
    for (uint8_t* target : targets_) {
        DeviceFree(target);
    }
    targets_.clear();
}

bool GaussianSplatting::EnsureImageCapacity(size_t pixels) {
    if (pixels <= image_pixels_) {
        return true;
    }
    float* grown = nullptr;
    if (!DeviceAllocateCount(&grown, pixels * 3)) {
        return false;
    }
    DeviceFree(image_);
    image_ = grown;
    image_pixels_ = pixels;
    return true;
}

bool GaussianSplatting::Initialize(const Scene& scene) {
    const auto& views = scene.train_cameras_;
    const auto& points = scene.point_cloud_;
    const bool usable = !gaussians_.gaussians_ && !views.empty() && views.size() == scene.train_images_.size() &&
                        !points.positions_.empty() && points.positions_.size() == points.colors_.size() &&
                        points.positions_.size() <= static_cast<size_t>(std::numeric_limits<int>::max()) &&
                        DefaultContext() != nullptr;
    if (!usable) {
        return false;
    }
    size_t largest = 0;
    for (size_t v = 0; v < views.size(); ++v) {
        if (!ImageMatches(scene.train_images_[v], views[v])) {
            GS_NATIVE_ERROR("training image %zu does not match its camera's %dx%d", v, views[v].image_size_.width,
                views[v].image_size_.height);
            return false;
        }
        largest = std::max(largest, scene.train_images_[v].data.size() / 3);
    }
    scene_ = scene;
    scene_extent_ = SceneExtent(scene_.train_cameras_, scene_.point_cloud_.positions_);
    ReleaseImages();
    if (!EnsureImageCapacity(largest)) {
        return false;
    }
    for (const Image8& image : scene_.train_images_) {
        uint8_t* resident = nullptr;
        bool ok = DeviceAllocateCount(&resident, image.data.size()) &&
                  gsn_memcpy_to_device(resident, image.data.data(), image.data.size()) == GSN_OK;
        if (!ok) {
            DeviceFree(resident);
            ReleaseImages();
            return false;
        }
        targets_.push_back(resident);
    }
    cycle_.resize(views.size());
    std::iota(cycle_.begin(), cycle_.end(), 0);
    cycle_position_ = cycle_.size();  // the first Train iteration shuffles
    completed_ = 0;
    sh_degree_ = 0;
    next_view_ = -1;

    const int count = static_cast<int>(points.positions_.size());
    std::vector<float> mean_distance2 = ComputeMeanDist2Cuda(scene_.point_cloud_.positions_);
    if (mean_distance2.size() != points.positions_.size()) {
        GS_NATIVE_ERROR("3-NN initialisation of %d points failed", count);
        return false;
    }
    std::vector<float> rows = InitialParameters(scene_.point_cloud_, mean_distance2);
    float* device_rows = nullptr;
    bool ok = DeviceAllocateCount(&device_rows, rows.size()) &&
              gsn_memcpy_to_device(device_rows, rows.data(), rows.size() * sizeof(float)) == GSN_OK &&
              density_.Initialize(count, count) && solver_.RemapGaussians(nullptr, count);
    if (!ok) {
        DeviceFree(device_rows);
        GS_NATIVE_ERROR("cannot place %d initial Gaussians on the device", count);
        return false;
    }
    gaussians_.gaussians_ = device_rows;
    gaussians_.gaussian_count_ = count;
    GS_NATIVE_INFO("initialised %d Gaussians from %zu views (extent %g)", count, views.size(), scene_extent_);
    return true;
}

int GaussianSplatting::NextCamera() {
    if (cycle_position_ >= cycle_.size()) {
        std::shuffle(cycle_.begin(), cycle_.end(), shuffler_);
        cycle_position_ = 0;
    }
    return cycle_[cycle_position_];
}

bool GaussianSplatting::Step() {
    const int view = NextCamera();
    const CameraParams camera = scene_.train_cameras_[static_cast<size_t>(view)].GetCameraParams();
    const int iteration = completed_ + 1;
    const schedule::Step plan = schedule::StepFor(iteration, scene_extent_);
    sh_degree_ = plan.sh_degree;
    // The solver's splat backward then prepares on a second stream beside the forward raster.
    gsn_context* context = DefaultContext();
    if (context && gsn_context_prefetch_backward(context) != GSN_OK) {
        GS_NATIVE_ERROR("arming the backward prefetch failed: %s", gsn_last_error());
        return false;
    }
    bool ok = true;
    if (next_view_ == view && next_count_ == gaussians_.gaussian_count_ && next_degree_ == sh_degree_) {
        // The previous gradient pass projected the current parameters for this camera.
        current_ = 1 - current_;
        ok = renderer().RenderProjected(gaussians_, camera, background_, image_);
    }
    else {
        ok = renderer().Render(gaussians_, camera, background_, image_, sh_degree_);
    }
    next_view_ = -1;
    const uint8_t* target = targets_[static_cast<size_t>(view)];
    // The next camera is known unless this pass over the views ends here (Train reseeds the
    // shuffle, so the next pass's order is drawn only when it starts).
    const bool fuse = cycle_position_ + 1 < cycle_.size() &&
                      schedule::StepFor(iteration + 1, scene_extent_).sh_degree == sh_degree_;
    if (ok && fuse) {
        const int next = cycle_[cycle_position_ + 1];
        const CameraParams next_camera = scene_.train_cameras_[static_cast<size_t>(next)].GetCameraParams();
        ok = solver_.IterateAndProject(camera, renderer(), image_, target, gaussians_, plan.position_learning_rate,
            sh_degree_, next_camera, renderers_[1 - current_]);
        if (ok) {
            next_view_ = next;
            next_count_ = gaussians_.gaussian_count_;
            next_degree_ = sh_degree_;
        }
    }
    else if (ok) {
        ok = solver_.Iterate(camera, renderer(), image_, target, gaussians_, plan.position_learning_rate, sh_degree_);
    }
    if (!ok) {
        GS_NATIVE_ERROR("training iteration %d on view %d stopped: %s", iteration, view, gsn_last_error());
        return false;
    }
    completed_ = iteration;
    ++cycle_position_;
    return Upkeep();
}

bool GaussianSplatting::Upkeep() {
    const schedule::Upkeep upkeep = schedule::UpkeepAfter(completed_);
    const int count = gaussians_.gaussian_count_;
    if (upkeep.accumulate) {
        const float* statistics = solver_.GetProjectionStatistics();
        const float* absolute = solver_.GetAbsoluteProjectionStatistics();
        if (!density_.AccumulateStatistics(statistics, absolute, count) ||
            !density_.AccumulateFeatures(statistics, absolute, renderer().GetDepths(),
                reinterpret_cast<const float*>(renderer().GetConicOpacities()), count)) {
            return false;
        }
    }
    if (upkeep.control_density || upkeep.opacity_cap > 0.0f) {
        next_view_ = -1;  // the parameters change: the projection made for the next camera is stale
    }
    if (upkeep.control_density) {
        int changes[2] = {0, 0};
        bool ok = density_.ControlDensity(gaussians_, scene_extent_, schedule::kGradientThreshold,
                      schedule::kPercentDense, schedule::kMinOpacity, upkeep.max_screen_size, 0, changes,
                      schedule::kAbsoluteGradientThreshold) &&
                  solver_.RemapGaussians(density_.GetHistoryParentIds(), gaussians_.gaussian_count_);
        if (!ok) {
            GS_NATIVE_ERROR("density step after iteration %d stopped: %s", completed_, gsn_last_error());
            return false;
        }
        GS_NATIVE_INFO("after iteration %d: %d -> %d Gaussians (+%d, -%d)", completed_, count,
            gaussians_.gaussian_count_, changes[0], changes[1]);
    }
    if (upkeep.opacity_cap > 0.0f) {
        if (!density_.ResetOpacity(gaussians_.gaussians_, gaussians_.gaussian_count_, upkeep.opacity_cap) ||
            !solver_.ResetOpacityHistory()) {
            return false;
        }
        GS_NATIVE_INFO("after iteration %d: opacity capped at %g", completed_, upkeep.opacity_cap);
    }
    return true;
}

bool GaussianSplatting::Train(int iterations) {
    if (iterations < 0 || !gaussians_.gaussians_ || targets_.empty()) {
        return false;
    }
    shuffler_.seed(0);
    for (int done = 0; done < iterations; ++done) {
        if (gaussians_.gaussian_count_ == 0 || !Step()) {
            return false;
        }
    }
    return Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize");
}

bool GaussianSplatting::Render(const Camera& source, ImageF& image) {
    const int width = source.image_size_.width;
    const int height = source.image_size_.height;
    if (!gaussians_.gaussians_ || width <= 0 || height <= 0) {
        return false;
    }
    const size_t pixels = static_cast<size_t>(width) * height;
    const CameraParams camera = source.GetCameraParams();
    if (!EnsureImageCapacity(pixels) || !renderer().Render(gaussians_, camera, background_, image_, sh_degree_) ||
        !Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize")) {
        return false;
    }
    image.width = width;
    image.height = height;
    image.data.resize(pixels * 3);
    return gsn_memcpy_to_host(image.data.data(), image_, pixels * 3 * sizeof(float)) == GSN_OK;
}

bool GaussianSplatting::Save(const std::string& path) {
    const int count = gaussians_.gaussian_count_;
    if (count <= 0 || !gaussians_.gaussians_) {
        return false;
    }
    std::vector<float> rows(static_cast<size_t>(count) * kMaxParamCount);
    if (!Check(gsn_context_synchronize(DefaultContext()), "gsn_context_synchronize") ||
        gsn_memcpy_to_host(rows.data(), gaussians_.gaussians_, rows.size() * sizeof(float)) != GSN_OK) {
        return false;
    }
    return WriteGaussiansPly(path, GaussiansFromRows(rows, count));
}

} // namespace gs_native
