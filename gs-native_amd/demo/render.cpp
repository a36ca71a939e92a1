// Renders the held-out (or training) views of a COLMAP scene with a trained Gaussian PLY and
// scores them: the command-line program of reference demo/render.cpp, over the library calls of
// include/gs_native/eval/evaluate.h.
//
//   gs_native_render --scene=<COLMAP dir> --model=<model.ply> --output=<dir>
//                    [--downscale=1] [--sh_degree=3] [--split=test|train]
//
// Writes <output>/<split>/gs-native/{renders,gt,comparison}/%05d.png and psnr.csv.
#include "eval/evaluate.h"
#include "gsn_hip.h"
#include "io/data_io.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

using namespace gs_native;

namespace {

bool ParseArguments(int argc, char** argv, std::map<std::string, std::string>& options) {
    for (int i = 1; i < argc; ++i) {
        const char* argument = argv[i];
        if (std::strncmp(argument, "--", 2) != 0 || !std::strchr(argument, '=')) {
            std::fprintf(stderr, "unexpected argument %s (use --name=value)\n", argument);
            return false;
        }
        const char* equals = std::strchr(argument, '=');
        options[std::string(argument + 2, equals)] = equals + 1;
    }
    return true;
}

} // namespace

int main(int argc, char** argv) {
    std::map<std::string, std::string> options = {{"downscale", "1"}, {"sh_degree", "3"}, {"split", "test"}};
    if (!ParseArguments(argc, argv, options) || !options.count("scene") || !options.count("model") ||
        !options.count("output") || (options["split"] != "test" && options["split"] != "train")) {
        std::fprintf(stderr, "usage: %s --scene=<dir> --model=<ply> --output=<dir> [--downscale=N] [--sh_degree=D] "
                             "[--split=test|train]\n", argv[0]);
        return 1;
    }
    const int downscale = std::max(1, std::atoi(options["downscale"].c_str()));
    const int sh_degree = std::clamp(std::atoi(options["sh_degree"].c_str()), 0, 3);
    Scene scene;
    std::vector<Gaussian> gaussians;
    if (!ReadColmapScene(options["scene"], downscale, scene) || !ReadGaussiansPly(options["model"], gaussians)) {
        return 1;
    }
    const bool test = options["split"] == "test";
    const std::vector<Camera>& cameras = test ? scene.test_cameras_ : scene.train_cameras_;
    const std::vector<Image8>& targets = test ? scene.test_images_ : scene.train_images_;
    GaussianBuffers buffers;
    if (cameras.empty() || !UploadGaussians(gaussians, buffers)) {
        std::fprintf(stderr, "no cameras or no valid Gaussians\n");
        return 1;
    }
    std::vector<ViewScore> scores;
    std::vector<Image8> renders;
    bool ok = EvaluateViews(buffers, cameras, targets, sh_degree, scores, &renders) &&
              WriteEvaluation(options["output"], options["split"], scores, renders, targets);
    gsn_device_free(buffers.gaussians_);
    if (!ok) {
        return 1;
    }
    for (size_t v = 0; v < scores.size(); ++v) {
        std::printf("View %zu (%s): PSNR %.6f dB\n", v, scores[v].image_name.c_str(), scores[v].psnr_db);
    }
    std::printf("Mean PSNR: %.6f dB\n", MeanPsnr(scores));
    return 0;
}
