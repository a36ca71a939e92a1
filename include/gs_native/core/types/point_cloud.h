// Sparse initial points (reference src/core/types/point_cloud.h).
#pragma once

#include "core/types/linalg.h"

#include <vector>

namespace gs_native {

struct PointCloud {
    std::vector<Vector3f> positions_;
    std::vector<Vector3f> colors_;
};

} // namespace gs_native
