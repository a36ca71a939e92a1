// Mean squared distance to the three nearest neighbours of every point (reference
// src/core/knn/knn.h:9). Accepts any point type with operator[] (e.g. gs_native::Vector3f or
// Eigen::Vector3f). Returns an empty vector on failure.
#pragma once

#include <cstddef>
#include <vector>

namespace gs_native {

std::vector<float> ComputeMeanDist2Flat(const float* xyz, size_t count);

template<typename Point>
std::vector<float> ComputeMeanDist2Cuda(const std::vector<Point>& points) {
    std::vector<float> xyz(points.size() * 3);
    for (size_t i = 0; i < points.size(); ++i) {
        for (int axis = 0; axis < 3; ++axis) {
            xyz[3 * i + axis] = static_cast<float>(points[i][axis]);
        }
    }
    return ComputeMeanDist2Flat(xyz.data(), points.size());
}

} // namespace gs_native
