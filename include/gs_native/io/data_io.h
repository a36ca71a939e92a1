// Scene and model files (reference src/io/data_io.h).
// Gaussian PLY (data_io.cpp:296-300, 420-590): the standard 62-float vertex layout (x y z,
// nx ny nz, f_dc_0..2, f_rest_0..44 channel-major, opacity, scale_0..2, rot_0..3 quaternion
// w x y z). Rotations are stored internally as axis-angle.
// COLMAP scenes (data_io.cpp:79-418): sparse/0/{cameras,images,points3D}.bin and images/ (PNG
// or binary PPM here; JPEG needs a decoder this build does not have).
#pragma once

#include "core/types/gaussian.h"
#include "core/types/scene.h"

#include <string>
#include <vector>

namespace gs_native {

bool WriteGaussiansPly(const std::string& path, const std::vector<Gaussian>& gaussians);

bool ReadGaussiansPly(const std::string& path, std::vector<Gaussian>& gaussians);

// Cameras, images (shrunk by `downscale` with area averaging when > 1), held-out split (every
// 8th image in name order) and the sparse point cloud of a COLMAP reconstruction.
bool ReadColmapScene(const std::string& scene_dir, int downscale, Scene& scene);

} // namespace gs_native
