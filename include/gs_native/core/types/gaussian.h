// Host-side Gaussian record (reference src/core/types/gaussian.h:10-42): position, log-scale,
// axis-angle rotation, opacity logit and 16x3 SH coefficients.
#pragma once

#include "core/types/linalg.h"

#include <algorithm>
#include <cmath>

namespace gs_native {

using SHCoeffs = MatrixF<16, 3>;

struct Gaussian {
    Vector3f position_ = Vector3f::Zero();
    Vector3f log_scale_ = Vector3f::Zero();
    Vector3f rotation_ = Vector3f::Zero();
    float opacity_logit_ = 0.0f;
    SHCoeffs sh_ = SHCoeffs::Zero();

    Vector3f GetScale() const {
        return MakeVector3f(std::exp(log_scale_[0]), std::exp(log_scale_[1]), std::exp(log_scale_[2]));
    }

    float GetMaxScale() const {
        return std::exp(std::max(log_scale_[0], std::max(log_scale_[1], log_scale_[2])));
    }

    Matrix3f GetRotationMatrix() const {
        float angle = rotation_.norm();
        if (angle < 1e-8f) {
            return Matrix3f::Identity();
        }
        float x = rotation_[0] / angle, y = rotation_[1] / angle, z = rotation_[2] / angle;
        float c = std::cos(angle), s = std::sin(angle), t = 1.0f - c;
        Matrix3f R;
        R.m = {c + t * x * x, t * x * y - s * z, t * x * z + s * y, t * x * y + s * z, c + t * y * y, t * y * z - s * x,
            t * x * z - s * y, t * y * z + s * x, c + t * z * z};
        return R;
    }

    float GetOpacity() const {
        return 1.0f / (1.0f + std::exp(-opacity_logit_));
    }
};

} // namespace gs_native
