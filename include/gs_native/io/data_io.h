// Gaussian PLY files (reference src/io/data_io.h, data_io.cpp:296-300, 420-590): the standard
// 62-float vertex layout (x y z, nx ny nz, f_dc_0..2, f_rest_0..44 channel-major, opacity,
// scale_0..2, rot_0..3 quaternion w x y z). Rotations are stored internally as axis-angle.
#pragma once

#include "core/types/gaussian.h"

#include <string>
#include <vector>

namespace gs_native {

bool WriteGaussiansPly(const std::string& path, const std::vector<Gaussian>& gaussians);

bool ReadGaussiansPly(const std::string& path, std::vector<Gaussian>& gaussians);

} // namespace gs_native
