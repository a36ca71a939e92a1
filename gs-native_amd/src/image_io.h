// 8-bit RGB image files and the area downscale of the COLMAP loader. OpenCV is not part of this
// build, so PNG (8-bit grey / grey+alpha / RGB / RGBA, non-interlaced; inflate and deflate by
// zlib) and binary PPM (P6, maxval 255) are read and written here. Other formats (JPEG) are
// rejected with an error message.
#pragma once

#include "core/types/scene.h"

#include <string>

namespace gs_native {

bool ReadImageRgb(const std::string& path, Image8& image);
bool WritePngRgb(const std::string& path, const Image8& image);

// Area-weighted downscale to width x height (each output pixel is the mean of the input area it
// covers, partial pixels weighted by their covered fraction), rounded half to even; the
// definition of OpenCV's INTER_AREA when shrinking.
Image8 DownscaleArea(const Image8& image, int width, int height);

} // namespace gs_native
