// Host-side camera (reference src/core/types/camera.h:12-41). K_, R_c_w_ are row-major 3x3,
// t_c_w_ a 3-vector; GetCameraParams() builds the device record (near 0.2, far FLT_MAX,
// 16x16 tiles).
#pragma once

#include "core/types/buffers.h"
#include "core/types/linalg.h"

#include <limits>
#include <string>

namespace gs_native {

struct ImageSize {
    int width = 0;
    int height = 0;
};

struct Camera {
    std::string image_name_;
    ImageSize image_size_;
    Matrix3f K_ = Matrix3f::Identity();
    Matrix3f R_c_w_ = Matrix3f::Identity();
    Vector3f t_c_w_ = Vector3f::Zero();
    float feature_depth_low_ = std::numeric_limits<float>::min();
    float feature_depth_high_ = std::numeric_limits<float>::max();

    CameraParams GetCameraParams() const {
        CameraParams params = {};
        params.width_ = image_size_.width;
        params.height_ = image_size_.height;
        params.fx_ = K_(0, 0);
        params.fy_ = K_(1, 1);
        params.cx_ = K_(0, 2);
        params.cy_ = K_(1, 2);
        for (int row = 0; row < 3; ++row) {
            params.t_c_w_[row] = t_c_w_[row];
            for (int col = 0; col < 3; ++col) {
                params.R_c_w_[row * 3 + col] = R_c_w_(row, col);
            }
        }
        params.near_depth_ = 0.2f;
        params.far_depth_ = std::numeric_limits<float>::max();
        params.tile_count_x_ = (params.width_ + 15) / 16;
        params.tile_count_y_ = (params.height_ + 15) / 16;
        return params;
    }
};

} // namespace gs_native
