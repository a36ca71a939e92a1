// Scene: cameras, their 8-bit RGB images (row-major HWC) and the initial point cloud
// (reference src/core/types/scene.h).
#pragma once

#include "core/types/camera.h"
#include "core/types/point_cloud.h"

#include <cstdint>
#include <vector>

namespace gs_native {

struct Image8 {
    int width = 0;
    int height = 0;
    std::vector<uint8_t> data;  // height * width * 3, RGB

    bool empty() const {
        return data.empty();
    }
};

struct ImageF {
    int width = 0;
    int height = 0;
    std::vector<float> data;  // height * width * 3, RGB, unclamped
};

struct Scene {
    std::vector<Camera> train_cameras_;
    std::vector<Camera> test_cameras_;
    std::vector<Image8> train_images_;
    std::vector<Image8> test_images_;
    PointCloud point_cloud_;
};

} // namespace gs_native
