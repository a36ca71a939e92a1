// Held-out evaluation of a trained model: the scoring and output of the reference's render program
// (demo/render.cpp:107-155) as library calls. Each view is rendered forward only (black
// background), clamped to [0, 1], quantised to 8 bits (x 255 in FP32, rounded half to even) and
// scored by PSNR against its 8-bit target; the report holds one line per view and the mean.
#pragma once

#include "core/types/buffers.h"
#include "core/types/gaussian.h"
#include "core/types/scene.h"

#include <string>
#include <vector>

namespace gs_native {

struct ViewScore {
    std::string image_name;
    double psnr_db = 0.0;
    double squared_error = 0.0;  // sum over pixels and channels of the 8-bit differences squared
};

// Scores every camera; `renders` (optional) receives the quantised images.
bool EvaluateViews(const GaussianBuffers& gaussians, const std::vector<Camera>& cameras,
    const std::vector<Image8>& targets, int sh_degree, std::vector<ViewScore>& scores,
    std::vector<Image8>* renders = nullptr);

// Mean PSNR over the views (the reference averages the per-view dB values).
double MeanPsnr(const std::vector<ViewScore>& scores);

// [58][N] device parameters of a Gaussian list (the layout of GaussianBuffers); the caller frees
// buffers.gaussians_ with gsn_device_free.
bool UploadGaussians(const std::vector<Gaussian>& gaussians, GaussianBuffers& buffers);

// <output>/<split>/gs-native/{renders,gt,comparison}/%05d.png and psnr.csv ("image,psnr_db", one
// line per view, then "mean,<dB>"). Refuses a non-empty destination, like the reference.
bool WriteEvaluation(const std::string& output_dir, const std::string& split, const std::vector<ViewScore>& scores,
    const std::vector<Image8>& renders, const std::vector<Image8>& targets);

} // namespace gs_native
