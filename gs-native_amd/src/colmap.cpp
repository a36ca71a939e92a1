// COLMAP scene loader (API of reference src/io/data_io.h; behaviour of data_io.cpp:79-418):
// sparse/0/{cameras,images,points3D}.bin in COLMAP's binary format, one PINHOLE camera, images
// from images/ sorted by name, optionally shrunk by an integer factor with area averaging, the
// intrinsics rescaled to the loaded size with the -0.5 pixel-centre shift, every 8th image (in
// name order, starting with the first) held out for testing, and per-view feature-depth ranges
// (1% / 99% quantiles of the depths of the view's observed points, at least 50 of them; views
// with fewer take the median of the others).
#include "io/data_io.h"

#include "image_io.h"
#include "log.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

namespace gs_native {

namespace {

constexpr int32_t kPinholeModel = 1;
constexpr int kHoldOutEvery = 8;
constexpr size_t kMinObservedPoints = 50;
constexpr float kLowDepthQuantile = 0.01f;
constexpr float kHighDepthQuantile = 0.99f;

// Little-endian binary reader that remembers the first failure.
class BinaryReader {
public:
    explicit BinaryReader(const std::string& path) : file_(path, std::ios::binary), path_(path) {}

    bool open() const {
        return static_cast<bool>(file_);
    }
    template<typename T>
    T Get() {
        T value{};
        file_.read(reinterpret_cast<char*>(&value), sizeof(T));
        return value;
    }
    void Skip(uint64_t bytes) {
        file_.ignore(static_cast<std::streamsize>(bytes));
    }
    std::string GetString() {
        std::string text;
        std::getline(file_, text, '\0');
        return text;
    }
    bool good() const {
        return static_cast<bool>(file_);
    }
    const std::string& path() const {
        return path_;
    }

private:
    std::ifstream file_;
    std::string path_;
};

struct PinholeCamera {
    int32_t id = 0;
    uint64_t width = 0, height = 0;
    double fx = 0, fy = 0, cx = 0, cy = 0;
};

struct RegisteredImage {
    int32_t camera_id = 0;
    std::string name;
    Matrix3f R_c_w = Matrix3f::Identity();
    Vector3f t_c_w = Vector3f::Zero();
    std::vector<int64_t> observed;  // point3D ids of the image's keypoints that have one
};

// Rotation of a unit quaternion (w, x, y, z), normalised first.
Matrix3f RotationOfQuaternion(float w, float x, float y, float z) {
    float length = std::sqrt(w * w + x * x + y * y + z * z);
    w /= length;
    x /= length;
    y /= length;
    z /= length;
    Matrix3f R;
    R(0, 0) = 1.0f - 2.0f * (y * y + z * z);
    R(0, 1) = 2.0f * (x * y - w * z);
    R(0, 2) = 2.0f * (x * z + w * y);
    R(1, 0) = 2.0f * (x * y + w * z);
    R(1, 1) = 1.0f - 2.0f * (x * x + z * z);
    R(1, 2) = 2.0f * (y * z - w * x);
    R(2, 0) = 2.0f * (x * z - w * y);
    R(2, 1) = 2.0f * (y * z + w * x);
    R(2, 2) = 1.0f - 2.0f * (x * x + y * y);
    return R;
}

bool ReadCameras(const std::string& path, PinholeCamera& camera) {
    BinaryReader in(path);
    if (!in.open()) {
        GS_NATIVE_ERROR("cannot open %s", path.c_str());
        return false;
    }
    uint64_t count = in.Get<uint64_t>();
    if (in.good() && count != 1) {
        GS_NATIVE_ERROR("%s holds %llu cameras; one shared PINHOLE camera is supported", path.c_str(),
            static_cast<unsigned long long>(count));
        return false;
    }
    camera.id = in.Get<int32_t>();
    int32_t model = in.Get<int32_t>();
    camera.width = in.Get<uint64_t>();
    camera.height = in.Get<uint64_t>();
    camera.fx = in.Get<double>();
    camera.fy = in.Get<double>();
    camera.cx = in.Get<double>();
    camera.cy = in.Get<double>();
    if (!in.good()) {
        GS_NATIVE_ERROR("%s is truncated", path.c_str());
        return false;
    }
    if (model != kPinholeModel || camera.width == 0 || camera.height == 0) {
        GS_NATIVE_ERROR("%s: camera model %d is not PINHOLE (1)", path.c_str(), model);
        return false;
    }
    return true;
}

bool ReadImages(const std::string& path, std::vector<RegisteredImage>& images) {
    BinaryReader in(path);
    if (!in.open()) {
        GS_NATIVE_ERROR("cannot open %s", path.c_str());
        return false;
    }
    uint64_t count = in.Get<uint64_t>();
    images.clear();
    for (uint64_t i = 0; in.good() && i < count; ++i) {
        RegisteredImage image;
        in.Get<int32_t>();  // image id (the order used is the name order)
        double q[4], t[3];
        for (double& value : q) {
            value = in.Get<double>();
        }
        for (double& value : t) {
            value = in.Get<double>();
        }
        image.R_c_w = RotationOfQuaternion(static_cast<float>(q[0]), static_cast<float>(q[1]),
            static_cast<float>(q[2]), static_cast<float>(q[3]));
        image.t_c_w = MakeVector3f(static_cast<float>(t[0]), static_cast<float>(t[1]), static_cast<float>(t[2]));
        image.camera_id = in.Get<int32_t>();
        image.name = in.GetString();
        uint64_t keypoints = in.Get<uint64_t>();
        for (uint64_t k = 0; in.good() && k < keypoints; ++k) {
            in.Skip(2 * sizeof(double));  // keypoint position
            int64_t point = in.Get<int64_t>();
            if (point >= 0) {
                image.observed.push_back(point);
            }
        }
        images.push_back(std::move(image));
    }
    if (!in.good()) {
        GS_NATIVE_ERROR("%s is truncated", path.c_str());
        return false;
    }
    return true;
}

bool ReadPoints(const std::string& path, PointCloud& cloud, std::unordered_map<int64_t, size_t>& index_of) {
    BinaryReader in(path);
    if (!in.open()) {
        GS_NATIVE_ERROR("cannot open %s", path.c_str());
        return false;
    }
    uint64_t count = in.Get<uint64_t>();
    cloud.positions_.clear();
    cloud.colors_.clear();
    index_of.clear();
    for (uint64_t i = 0; in.good() && i < count; ++i) {
        uint64_t id = in.Get<uint64_t>();
        double xyz[3];
        for (double& value : xyz) {
            value = in.Get<double>();
        }
        uint8_t rgb[3];
        for (uint8_t& value : rgb) {
            value = in.Get<uint8_t>();
        }
        in.Get<double>();                         // reprojection error
        in.Skip(in.Get<uint64_t>() * 2 * sizeof(int32_t));  // track: (image id, keypoint index) pairs
        if (!in.good()) {
            break;
        }
        index_of[static_cast<int64_t>(id)] = cloud.positions_.size();
        cloud.positions_.push_back(
            MakeVector3f(static_cast<float>(xyz[0]), static_cast<float>(xyz[1]), static_cast<float>(xyz[2])));
        cloud.colors_.push_back(MakeVector3f(rgb[0] / 255.0f, rgb[1] / 255.0f, rgb[2] / 255.0f));
    }
    if (!in.good()) {
        GS_NATIVE_ERROR("%s is truncated", path.c_str());
        return false;
    }
    return true;
}

// 1% and 99% depth quantiles of the view's observed points in front of it; false with fewer
// than 50 such points.
bool DepthRange(const RegisteredImage& image, const PointCloud& cloud,
    const std::unordered_map<int64_t, size_t>& index_of, Camera& camera) {
    std::vector<float> depths;
    for (int64_t point : image.observed) {
        auto found = index_of.find(point);
        if (found == index_of.end()) {
            continue;
        }
        const Vector3f& p = cloud.positions_[found->second];
        float depth = image.R_c_w(2, 0) * p[0] + image.R_c_w(2, 1) * p[1] + image.R_c_w(2, 2) * p[2] + image.t_c_w[2];
        if (depth > 0.0f) {
            depths.push_back(depth);
        }
    }
    if (depths.size() < kMinObservedPoints) {
        return false;
    }
    std::sort(depths.begin(), depths.end());
    const float last = static_cast<float>(depths.size() - 1);
    camera.feature_depth_low_ = depths[static_cast<size_t>(kLowDepthQuantile * last)];
    camera.feature_depth_high_ = depths[static_cast<size_t>(kHighDepthQuantile * last)];
    return true;
}

float UpperMedian(std::vector<float> values) {
    std::nth_element(values.begin(), values.begin() + static_cast<std::ptrdiff_t>(values.size() / 2), values.end());
    return values[values.size() / 2];
}

} // namespace

bool ReadColmapScene(const std::string& scene_dir, int downscale, Scene& scene) {
    const std::string sparse = scene_dir + "/sparse/0/";
    PinholeCamera intrinsics;
    std::vector<RegisteredImage> images;
    PointCloud cloud;
    std::unordered_map<int64_t, size_t> index_of;
    if (!ReadCameras(sparse + "cameras.bin", intrinsics) || !ReadImages(sparse + "images.bin", images) ||
        !ReadPoints(sparse + "points3D.bin", cloud, index_of)) {
        return false;
    }
    std::sort(images.begin(), images.end(),
        [](const RegisteredImage& a, const RegisteredImage& b) { return a.name < b.name; });
    Scene loaded;
    loaded.point_cloud_ = std::move(cloud);
    std::vector<bool> train_ranged, test_ranged;
    for (size_t i = 0; i < images.size(); ++i) {
        const RegisteredImage& registered = images[i];
        if (registered.camera_id != intrinsics.id) {
            GS_NATIVE_ERROR("image %s refers to camera %d, not %d", registered.name.c_str(), registered.camera_id,
                intrinsics.id);
            return false;
        }
        Image8 pixels;
        if (!ReadImageRgb(scene_dir + "/images/" + registered.name, pixels)) {
            return false;
        }
        if (downscale > 1) {
            int width = static_cast<int>(std::lround(pixels.width / static_cast<float>(downscale)));
            int height = static_cast<int>(std::lround(pixels.height / static_cast<float>(downscale)));
            pixels = DownscaleArea(pixels, std::max(width, 1), std::max(height, 1));
        }
        Camera camera;
        camera.image_name_ = registered.name;
        camera.image_size_ = {pixels.width, pixels.height};
        const float sx = pixels.width / static_cast<float>(intrinsics.width);
        const float sy = pixels.height / static_cast<float>(intrinsics.height);
        camera.K_ = Matrix3f::Identity();
        camera.K_(0, 0) = static_cast<float>(intrinsics.fx) * sx;
        camera.K_(1, 1) = static_cast<float>(intrinsics.fy) * sy;
        camera.K_(0, 2) = static_cast<float>(intrinsics.cx) * sx - 0.5f;
        camera.K_(1, 2) = static_cast<float>(intrinsics.cy) * sy - 0.5f;
        camera.R_c_w_ = registered.R_c_w;
        camera.t_c_w_ = registered.t_c_w;
        const bool ranged = DepthRange(registered, loaded.point_cloud_, index_of, camera);
        if (i % kHoldOutEvery == 0) {
            loaded.test_cameras_.push_back(camera);
            loaded.test_images_.push_back(std::move(pixels));
            test_ranged.push_back(ranged);
        }
        else {
            loaded.train_cameras_.push_back(camera);
            loaded.train_images_.push_back(std::move(pixels));
            train_ranged.push_back(ranged);
        }
    }
    // Views without enough observed points take the median range of the others.
    std::vector<float> lows, highs;
    auto collect = [&](const std::vector<Camera>& cameras, const std::vector<bool>& ranged) {
        for (size_t i = 0; i < cameras.size(); ++i) {
            if (ranged[i]) {
                lows.push_back(cameras[i].feature_depth_low_);
                highs.push_back(cameras[i].feature_depth_high_);
            }
        }
    };
    collect(loaded.train_cameras_, train_ranged);
    collect(loaded.test_cameras_, test_ranged);
    if (!lows.empty()) {
        const float low = UpperMedian(lows), high = UpperMedian(highs);
        auto fill = [&](std::vector<Camera>& cameras, const std::vector<bool>& ranged) {
            for (size_t i = 0; i < cameras.size(); ++i) {
                if (!ranged[i]) {
                    cameras[i].feature_depth_low_ = low;
                    cameras[i].feature_depth_high_ = high;
                }
            }
        };
        fill(loaded.train_cameras_, train_ranged);
        fill(loaded.test_cameras_, test_ranged);
    }
    scene = std::move(loaded);
    GS_NATIVE_INFO("loaded %s: %zu training views, %zu held-out views, %zu points", scene_dir.c_str(),
        scene.train_cameras_.size(), scene.test_cameras_.size(), scene.point_cloud_.positions_.size());
    return true;
}

} // namespace gs_native
