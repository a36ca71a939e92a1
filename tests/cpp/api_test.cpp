// C++ API tests of the gs_native drop-in classes, following the reference's integration and
// class-level cases (test/gaussian_splatting_test.cpp, density_test.cu, knn_test.cu,
// solver_test.cu:204-236). Built by __graft_entry__.build(); run by tests/test_gpu_cpp_api.py.
#include "core/density/density_controller.h"
#include "core/gaussian_splatting.h"
#include "core/knn/knn.h"
#include "core/render/renderer.h"
#include "core/solver/solver.h"
#include "io/data_io.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>
#include <unistd.h>

using namespace gs_native;

namespace {

int g_failures = 0;

#define EXPECT(condition)                                                             \
    do {                                                                              \
        if (!(condition)) {                                                           \
            std::printf("  expectation failed: %s (%s:%d)\n", #condition, __FILE__, __LINE__); \
            ++g_failures;                                                             \
            return;                                                                   \
        }                                                                             \
    } while (0)

template<typename T>
struct DeviceVector {
    T* data = nullptr;
    size_t count = 0;
    explicit DeviceVector(const std::vector<T>& values) : count(values.size()) {
        if (count) {
            if (hipMalloc(reinterpret_cast<void**>(&data), count * sizeof(T)) != hipSuccess ||
                hipMemcpy(data, values.data(), count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
                std::abort();
            }
        }
    }
    explicit DeviceVector(size_t n) : count(n) {
        if (count && hipMalloc(reinterpret_cast<void**>(&data), count * sizeof(T)) != hipSuccess) {
            std::abort();
        }
    }
    ~DeviceVector() {
        if (data) {
            (void)hipFree(data);
        }
    }
    std::vector<T> Download() const {
        std::vector<T> out(count);
        if (count && hipMemcpy(out.data(), data, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
            std::abort();
        }
        return out;
    }
};

template<typename T>
std::vector<T> ReadDevice(const T* pointer, size_t count) {
    std::vector<T> out(count);
    if (count && hipMemcpy(out.data(), pointer, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) {
        std::abort();
    }
    return out;
}

// test_support.h:113-130
std::vector<Gaussian> MakeGaussians(int count) {
    std::vector<Gaussian> gaussians(count);
    for (int i = 0; i < count; ++i) {
        Gaussian& g = gaussians[i];
        g.position_ = MakeVector3f((i % 16 - 7.5f) * 0.035f, ((i / 16) % 16 - 7.5f) * 0.035f, 3.0f + 0.001f * i);
        g.log_scale_ = MakeVector3f(std::log(0.12f), std::log(0.16f), std::log(0.2f));
        g.rotation_ = MakeVector3f(0.15f, -0.2f, 0.1f);
        g.opacity_logit_ = -1.0f;
        for (int c = 0; c < 3; ++c) {
            g.sh_(0, c) = (0.6f - 0.1f * c - 0.5f) / 0.28209479177387814f;
            for (int k = 1; k < 16; ++k) {
                g.sh_(k, c) = 0.02f * std::sin(float(k + c + i));
            }
        }
    }
    return gaussians;
}

std::vector<float> PackGaussians(const std::vector<Gaussian>& gaussians) {
    size_t count = gaussians.size();
    std::vector<float> values(count * kMaxParamCount);
    for (size_t i = 0; i < count; ++i) {
        for (int axis = 0; axis < 3; ++axis) {
            values[axis * count + i] = gaussians[i].position_[axis];
            values[(kLogScaleIndex + axis) * count + i] = gaussians[i].log_scale_[axis];
            values[(kRotationIndex + axis) * count + i] = gaussians[i].rotation_[axis];
            for (int k = 0; k < 16; ++k) {
                values[(kShIndex + axis * 16 + k) * count + i] = gaussians[i].sh_(k, axis);
            }
        }
        values[kOpacityIndex * count + i] = gaussians[i].opacity_logit_;
    }
    return values;
}

// test_support.h:132-148
CameraParams MakeCamera(int width, int height) {
    CameraParams camera = {};
    camera.width_ = width;
    camera.height_ = height;
    camera.fx_ = 20.0f;
    camera.fy_ = 21.0f;
    camera.cx_ = 0.5f * (width - 1);
    camera.cy_ = 0.5f * (height - 1);
    camera.R_c_w_[0] = camera.R_c_w_[4] = camera.R_c_w_[8] = 1;
    camera.near_depth_ = 0.2f;
    camera.far_depth_ = 100.0f;
    camera.tile_count_x_ = (width + 15) / 16;
    camera.tile_count_y_ = (height + 15) / 16;
    return camera;
}

// test_support.h:150-162
Scene MakeScene() {
    Scene scene;
    Camera camera;
    camera.image_name_ = "synthetic.png";
    camera.image_size_ = {9, 7};
    camera.K_.m = {20, 0, 4, 0, 21, 3, 0, 0, 1};
    scene.train_cameras_.push_back(camera);
    Image8 image;
    image.width = 9;
    image.height = 7;
    image.data.resize(9 * 7 * 3);
    for (size_t p = 0; p < image.data.size(); p += 3) {
        image.data[p] = 160;
        image.data[p + 1] = 170;
        image.data[p + 2] = 180;
    }
    scene.train_images_.push_back(image);
    scene.point_cloud_.positions_ = {MakeVector3f(-0.1f, -0.1f, 3), MakeVector3f(0.1f, -0.1f, 3),
        MakeVector3f(-0.1f, 0.1f, 3), MakeVector3f(0.1f, 0.1f, 3)};
    scene.point_cloud_.colors_.assign(4, MakeVector3f(0.6f, 0.5f, 0.4f));
    return scene;
}

// cpu_reference.h:120-152 (0.8 L1 + 0.2 (1 - SSIM), 11-tap sigma 1.5, zero border).
double ComputeLoss(const ImageF& image, const Image8& target) {
    int w = image.width, h = image.height;
    double kernel[11], sum = 0;
    for (int i = 0; i < 11; ++i) {
        kernel[i] = std::exp(-(i - 5.0) * (i - 5.0) / 4.5);
        sum += kernel[i];
    }
    for (double& k : kernel) {
        k /= sum;
    }
    auto filter = [&](const std::vector<double>& in) {
        std::vector<double> tmp(in.size()), out(in.size());
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                double s = 0;
                for (int k = 0; k < 11; ++k) {
                    int xx = x + k - 5;
                    if (xx >= 0 && xx < w) s += kernel[k] * in[y * w + xx];
                }
                tmp[y * w + x] = s;
            }
        }
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                double s = 0;
                for (int k = 0; k < 11; ++k) {
                    int yy = y + k - 5;
                    if (yy >= 0 && yy < h) s += kernel[k] * tmp[yy * w + x];
                }
                out[y * w + x] = s;
            }
        }
        return out;
    };
    double l1 = 0, ssim = 0;
    for (int c = 0; c < 3; ++c) {
        std::vector<double> a(w * h), b(w * h), aa(w * h), bb(w * h), ab(w * h);
        for (int p = 0; p < w * h; ++p) {
            a[p] = image.data[p * 3 + c];
            b[p] = target.data[p * 3 + c] / 255.0;
            aa[p] = a[p] * a[p];
            bb[p] = b[p] * b[p];
            ab[p] = a[p] * b[p];
            l1 += std::abs(a[p] - b[p]);
        }
        auto ma = filter(a), mb = filter(b), faa = filter(aa), fbb = filter(bb), fab = filter(ab);
        for (int p = 0; p < w * h; ++p) {
            double va = faa[p] - ma[p] * ma[p], vb = fbb[p] - mb[p] * mb[p], cov = fab[p] - ma[p] * mb[p];
            ssim += (2 * ma[p] * mb[p] + 1e-4) * (2 * cov + 9e-4) / ((ma[p] * ma[p] + mb[p] * mb[p] + 1e-4) * (va + vb + 9e-4));
        }
    }
    double count = 3.0 * w * h;
    return 0.8 * l1 / count + 0.2 * (1 - ssim / count);
}

bool AllFinite(const ImageF& image) {
    return std::all_of(image.data.begin(), image.data.end(), [](float v) { return std::isfinite(v); });
}

std::string TempPath(const char* name) {
    return std::string("/tmp/gs_native_api_test_") + std::to_string(getpid()) + "_" + name;
}

// gaussian_splatting_test.cpp:10-37
void InitializesIteratesSavesAndReloads() {
    Scene scene = MakeScene();
    GaussianSplatting model;
    EXPECT(model.Initialize(scene));
    ImageF before, after;
    EXPECT(model.Render(scene.train_cameras_[0], before));
    EXPECT(model.Train(3));
    EXPECT(model.Render(scene.train_cameras_[0], after));
    EXPECT(AllFinite(before) && AllFinite(after));
    EXPECT(ComputeLoss(after, scene.train_images_[0]) < ComputeLoss(before, scene.train_images_[0]));
    std::string path = TempPath("model.ply");
    EXPECT(model.Save(path));
    std::vector<Gaussian> loaded;
    EXPECT(ReadGaussiansPly(path, loaded));
    std::remove(path.c_str());
    EXPECT(loaded.size() == 4u);
    DeviceVector<float> parameters(PackGaussians(loaded));
    GaussianBuffers buffers{parameters.data, 4};
    CameraParams camera = MakeCamera(9, 7);
    Renderer renderer;
    DeviceVector<float> image(9 * 7 * 3);
    float background[3] = {};
    EXPECT(renderer.Render(buffers, camera, background, image.data, 0));
    std::vector<float> reloaded = image.Download();
    float worst = 0;
    for (size_t i = 0; i < reloaded.size(); ++i) {
        worst = std::max(worst, std::abs(reloaded[i] - after.data[i]));
    }
    EXPECT(worst < 2e-5f);
}

// gaussian_splatting_test.cpp:39-58
void CrossesDensityControlBoundary() {
    Scene scene = MakeScene();
    GaussianSplatting model;
    EXPECT(model.Initialize(scene));
    EXPECT(model.Train(600));
    ImageF image;
    EXPECT(model.Render(scene.train_cameras_[0], image));
    EXPECT(AllFinite(image));
    std::string path = TempPath("density.ply");
    EXPECT(model.Save(path));
    std::vector<Gaussian> gaussians;
    EXPECT(ReadGaussiansPly(path, gaussians));
    std::remove(path.c_str());
    std::printf("  gaussians after 600 iterations: %zu\n", gaussians.size());
    EXPECT(gaussians.size() == 2 * scene.point_cloud_.positions_.size());
    for (const Gaussian& g : gaussians) {
        EXPECT(g.position_.allFinite() && g.log_scale_.allFinite() && g.sh_.allFinite());
    }
}

// gaussian_splatting_test.cpp:60-72
void RejectsEmptySceneAndSupportsZeroIterations() {
    GaussianSplatting model;
    EXPECT(!model.Initialize(Scene{}));
    Scene scene = MakeScene();
    EXPECT(model.Initialize(scene));
    ImageF before, after;
    EXPECT(model.Render(scene.train_cameras_[0], before));
    EXPECT(model.Train(0));
    EXPECT(model.Render(scene.train_cameras_[0], after));
    EXPECT(before.data == after.data);
}

// density_test.cu:55-73 and 196-209 through the class.
void DensityControllerPrunesAndSwaps() {
    auto gaussians = MakeGaussians(3);
    gaussians[1].opacity_logit_ = std::log(0.001f / 0.999f);
    DeviceVector<float> original(PackGaussians(gaussians));
    DensityController controller;
    EXPECT(controller.Initialize(3, 3));
    float* copy = nullptr;
    EXPECT(hipMalloc(reinterpret_cast<void**>(&copy), original.count * sizeof(float)) == hipSuccess);
    EXPECT(hipMemcpy(copy, original.data, original.count * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess);
    GaussianBuffers buffers{copy, 3};
    int changes[2] = {};
    EXPECT(controller.ControlDensity(buffers, 1.0f, 0.15f, 0.01f, 0.005f, 0.0f, 42, changes));
    EXPECT(buffers.gaussian_count_ == 2);
    EXPECT(ReadDevice(controller.GetHistoryParentIds(), 2) == (std::vector<int>{0, 2}));
    EXPECT(changes[0] == 0 && changes[1] == 1);
    EXPECT(ReadDevice(controller.GetGaussianCount(), 1)[0] == 2);
    auto values = ReadDevice(buffers.gaussians_, 2 * kMaxParamCount);
    auto packed = PackGaussians(gaussians);
    for (int field = 0; field < kMaxParamCount; ++field) {
        EXPECT(values[field * 2] == packed[field * 3] && values[field * 2 + 1] == packed[field * 3 + 2]);
    }
    (void)hipFree(buffers.gaussians_);  // the caller owns the swapped-in buffer

    auto pruned = MakeGaussians(2);
    for (auto& g : pruned) g.opacity_logit_ = -12;
    DeviceVector<float> all(PackGaussians(pruned));
    DensityController second;
    EXPECT(second.Initialize(2, 2));
    float* buffer = nullptr;
    EXPECT(hipMalloc(reinterpret_cast<void**>(&buffer), all.count * sizeof(float)) == hipSuccess);
    EXPECT(hipMemcpy(buffer, all.data, all.count * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess);
    GaussianBuffers everything{buffer, 2};
    EXPECT(second.ControlDensity(everything, 1.0f, 0.15f, 0.01f, 0.005f, 0.0f, 42));
    EXPECT(everything.gaussian_count_ == 0 && everything.gaussians_ == nullptr);
}

// Learned densification hook (no reference counterpart): a proposal that always answers keep adds
// nothing, one that always answers clone clones every survivor; pruning stays the rule's.
void ProposalNetworkDrivesDensityControl() {
    int n = ProposalNetwork::ParameterCount();
    EXPECT(n == 5443);
    std::vector<float> keep(static_cast<size_t>(n), 0.0f);
    keep[n - 3] = 1.0f;
    ProposalNetwork keep_all;
    EXPECT(keep_all.Load(keep));
    EXPECT(keep_all.Parameters() == keep);
    DeviceVector<float> features(std::vector<float>(5 * 16, 0.5f));
    DeviceVector<float> logits(5 * 3);
    EXPECT(keep_all.Predict(features.data, 5, logits.data));
    auto values = logits.Download();
    for (int i = 0; i < 5; ++i) {
        EXPECT(values[3 * i] == 1.0f && values[3 * i + 1] == 0.0f && values[3 * i + 2] == 0.0f);
    }
    std::vector<float> clone(static_cast<size_t>(n), 0.0f);
    clone[n - 2] = 1.0f;
    ProposalNetwork clone_all;
    EXPECT(clone_all.Load(clone));

    for (int variant = 0; variant < 2; ++variant) {
        auto gaussians = MakeGaussians(3);
        gaussians[1].opacity_logit_ = std::log(0.001f / 0.999f);
        DeviceVector<float> original(PackGaussians(gaussians));
        DensityController controller;
        EXPECT(controller.Initialize(3, 3));
        controller.SetProposal(variant == 0 ? &keep_all : &clone_all);
        // every Gaussian observed once with a screen gradient far above the threshold
        DeviceVector<float> statistics(std::vector<float>{1.0f, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f});
        DeviceVector<float> absolute(std::vector<float>{1.0f, 1.0f, 1.0f, 0.0f, 0.0f, 0.0f});
        EXPECT(controller.AccumulateStatistics(statistics.data, absolute.data, 3));
        float* copy = nullptr;
        EXPECT(hipMalloc(reinterpret_cast<void**>(&copy), original.count * sizeof(float)) == hipSuccess);
        EXPECT(hipMemcpy(copy, original.data, original.count * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess);
        GaussianBuffers buffers{copy, 3};
        int changes[2] = {};
        EXPECT(controller.ControlDensity(buffers, 1.0f, 0.15f, 0.01f, 0.005f, 0.0f, 42, changes));
        if (variant == 0) {
            EXPECT(changes[0] == 0 && changes[1] == 1 && buffers.gaussian_count_ == 2);
        }
        else {
            EXPECT(changes[0] == 2 && changes[1] == 1 && buffers.gaussian_count_ == 4);
        }
        EXPECT(controller.GetProposalFallbacks() == 0);
        (void)hipFree(buffers.gaussians_);
    }

    // Growth guard: with gradients below the threshold the rule adds nothing, so a network that
    // splits everything is not followed and the step matches the rule (one prune, no additions).
    std::vector<float> split(static_cast<size_t>(n), 0.0f);
    split[n - 1] = 1.0f;
    ProposalNetwork split_all;
    EXPECT(split_all.Load(split));
    auto gaussians = MakeGaussians(3);
    gaussians[1].opacity_logit_ = std::log(0.001f / 0.999f);
    DeviceVector<float> original(PackGaussians(gaussians));
    DensityController controller;
    EXPECT(controller.Initialize(3, 3));
    controller.SetProposal(&split_all);
    DeviceVector<float> statistics(std::vector<float>{0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 1.0f});
    DeviceVector<float> absolute(std::vector<float>{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f});
    EXPECT(controller.AccumulateStatistics(statistics.data, absolute.data, 3));
    float* copy = nullptr;
    EXPECT(hipMalloc(reinterpret_cast<void**>(&copy), original.count * sizeof(float)) == hipSuccess);
    EXPECT(hipMemcpy(copy, original.data, original.count * sizeof(float), hipMemcpyDeviceToDevice) == hipSuccess);
    GaussianBuffers buffers{copy, 3};
    int changes[2] = {};
    EXPECT(controller.ControlDensity(buffers, 1.0f, 0.15f, 0.01f, 0.005f, 0.0f, 42, changes));
    EXPECT(changes[0] == 0 && changes[1] == 1 && buffers.gaussian_count_ == 2);
    EXPECT(controller.GetProposalFallbacks() == 1);
    (void)hipFree(buffers.gaussians_);
}

// solver_test.cu:320-336: opacity history reset -> first-step magnitude at step 2.
void SolverOpacityHistoryReset() {
    CameraParams camera = MakeCamera(9, 7);
    auto gaussians = MakeGaussians(1);
    gaussians[0].position_ = MakeVector3f(0.0f, 0.0f, 3.0f);
    DeviceVector<float> parameters(PackGaussians(gaussians));
    GaussianBuffers buffers{parameters.data, 1};
    DeviceVector<uint8_t> target(std::vector<uint8_t>(9 * 7 * 3, 200));
    DeviceVector<float> image(9 * 7 * 3);
    Renderer renderer;
    Solver solver;
    float background[3] = {};
    EXPECT(renderer.Render(buffers, camera, background, image.data, 0));
    EXPECT(solver.Iterate(camera, renderer, image.data, target.data, buffers, 0.0001f, 0));
    auto values = parameters.Download();
    values[kOpacityIndex] = std::log(0.01f / 0.99f);
    EXPECT(hipMemcpy(parameters.data, values.data(), values.size() * sizeof(float), hipMemcpyHostToDevice) == hipSuccess);
    EXPECT(solver.ResetOpacityHistory());
    EXPECT(renderer.Render(buffers, camera, background, image.data, 0));
    EXPECT(solver.Iterate(camera, renderer, image.data, target.data, buffers, 0.0001f, 0));
    auto actual = parameters.Download();
    double factor = (0.1 / (1 - std::pow(0.9, 2))) / std::sqrt(0.001 / (1 - std::pow(0.999, 2)));
    EXPECT(std::abs(std::abs(actual[kOpacityIndex] - values[kOpacityIndex]) - 0.025 * factor) < 2e-6);
}

// Solver::IterateAndProject (the next view's projection fused into the gradient pass) against
// Iterate followed by Render of that view: parameters and the rendered image bit for bit, at SH
// degree 3 on 300 Gaussians.
void SolverIterateAndProjectMatchesIterateThenRender() {
    CameraParams camera = MakeCamera(40, 36);
    CameraParams next_camera = MakeCamera(40, 36);
    next_camera.t_c_w_[0] = 0.05f;
    next_camera.fx_ = 22.0f;
    auto gaussians = MakeGaussians(300);
    DeviceVector<uint8_t> target(std::vector<uint8_t>(40 * 36 * 3, 150));
    float background[3] = {0.1f, 0.2f, 0.3f};
    std::vector<float> results[2];
    std::vector<float> images[2];
    for (int fused = 0; fused < 2; ++fused) {
        DeviceVector<float> parameters(PackGaussians(gaussians));
        GaussianBuffers buffers{parameters.data, 300};
        DeviceVector<float> image(40 * 36 * 3);
        DeviceVector<float> next_image(40 * 36 * 3);
        Renderer renderer;
        Renderer next_renderer;
        Solver solver;
        EXPECT(renderer.Render(buffers, camera, background, image.data, 3));
        if (fused) {
            EXPECT(!solver.IterateAndProject(camera, renderer, image.data, target.data, buffers, 0.0001f, 3, next_camera,
                renderer));
            EXPECT(solver.IterateAndProject(camera, renderer, image.data, target.data, buffers, 0.0001f, 3, next_camera,
                next_renderer));
            EXPECT(next_renderer.RenderProjected(buffers, next_camera, background, next_image.data));
        }
        else {
            EXPECT(solver.Iterate(camera, renderer, image.data, target.data, buffers, 0.0001f, 3));
            EXPECT(next_renderer.Render(buffers, next_camera, background, next_image.data, 3));
        }
        results[fused] = parameters.Download();
        images[fused] = next_image.Download();
    }
    EXPECT(std::memcmp(results[0].data(), results[1].data(), results[0].size() * sizeof(float)) == 0);
    EXPECT(std::memcmp(images[0].data(), images[1].data(), images[0].size() * sizeof(float)) == 0);
    double total = 0.0;
    for (float v : images[1]) {
        total += v;
    }
    EXPECT(total > 0.0);
}

// knn_test.cu:44-53 through the templated host API.
void KnnPreservesOrder() {
    std::vector<Vector3f> points = {MakeVector3f(100, 0, 0), MakeVector3f(0, 0, 0), MakeVector3f(1, 0, 0),
        MakeVector3f(3, 0, 0), MakeVector3f(7, 0, 0)};
    auto first = ComputeMeanDist2Cuda(points);
    EXPECT(first.size() == points.size());
    std::reverse(points.begin(), points.end());
    auto second = ComputeMeanDist2Cuda(points);
    std::reverse(second.begin(), second.end());
    EXPECT(first == second);
    EXPECT(std::abs(first[1] - (1.0f + 9.0f + 49.0f) / 3.0f) < 1e-6f);
    EXPECT(ComputeMeanDist2Cuda(std::vector<Vector3f>{}).empty());
    EXPECT(ComputeMeanDist2Cuda(std::vector<Vector3f>{MakeVector3f(1, 2, 3)}) == std::vector<float>{0.0f});
}

} // namespace

int main(int argc, char** argv) {
    std::vector<std::pair<std::string, std::function<void()>>> tests = {
        {"InitializesIteratesSavesAndReloads", InitializesIteratesSavesAndReloads},
        {"SolverIterateAndProjectMatchesIterateThenRender", SolverIterateAndProjectMatchesIterateThenRender},
        {"CrossesDensityControlBoundary", CrossesDensityControlBoundary},
        {"RejectsEmptySceneAndSupportsZeroIterations", RejectsEmptySceneAndSupportsZeroIterations},
        {"DensityControllerPrunesAndSwaps", DensityControllerPrunesAndSwaps},
        {"SolverOpacityHistoryReset", SolverOpacityHistoryReset},
        {"KnnPreservesOrder", KnnPreservesOrder},
        {"ProposalNetworkDrivesDensityControl", ProposalNetworkDrivesDensityControl},
    };
    int failed = 0;
    for (auto& [name, test] : tests) {
        if (argc > 1 && name != argv[1]) {
            continue;
        }
        std::printf("[ RUN  ] %s\n", name.c_str());
        std::fflush(stdout);
        int before = g_failures;
        test();
        bool ok = g_failures == before;
        failed += ok ? 0 : 1;
        std::printf("[ %s ] %s\n", ok ? " OK " : "FAIL", name.c_str());
        std::fflush(stdout);
    }
    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
