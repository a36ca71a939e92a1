// CPU checks of the file formats (no GPU): the COLMAP loader on the fixture of the reference's
// test/io_test.cpp:20-70 (written by tests/test_io.py with an independent PNG encoder), the PNG /
// PPM decoders and the area downscale against expectations written by the same script, and the
// Gaussian PLY round trip. Usage: io_test <fixture directory>. Prints one line per check.
#include "io/data_io.h"

#include "../../gs-native_amd/src/image_io.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

using namespace gs_native;

namespace {

int failures = 0;

void Expect(bool condition, const char* what) {
    std::printf("%s %s\n", condition ? "ok" : "FAILED", what);
    if (!condition) {
        ++failures;
    }
}

std::vector<uint8_t> ReadBytes(const std::string& path) {
    std::ifstream file(path, std::ios::binary);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(file), std::istreambuf_iterator<char>());
}

void ColmapFixture(const std::string& root) {
    Scene scene;
    Expect(ReadColmapScene(root + "/colmap", 2, scene), "colmap: the fixture loads");
    Expect(scene.train_cameras_.size() == 7 && scene.test_cameras_.size() == 2, "colmap: 7 training / 2 held-out views");
    Expect(scene.train_images_.size() == 7 && scene.test_images_.size() == 2, "colmap: one image per view");
    if (scene.test_cameras_.size() != 2 || scene.train_cameras_.empty()) {
        return;
    }
    Expect(scene.test_cameras_[0].image_name_ == "frame00.png" && scene.test_cameras_[1].image_name_ == "frame08.png",
        "colmap: held-out views are every 8th in name order");
    const Camera& camera = scene.train_cameras_[0];
    Expect(camera.image_size_.width == 10 && camera.image_size_.height == 9, "colmap: 19x17 shrunk by 2 is 10x9");
    Expect(std::fabs(camera.K_(0, 0) - 20.0 * 10 / 19) < 1e-5 && std::fabs(camera.K_(1, 1) - 21.0 * 9 / 17) < 1e-5,
        "colmap: focal lengths scaled to the loaded size");
    Expect(std::fabs(camera.K_(0, 2) - (9.0 * 10 / 19 - 0.5)) < 1e-5 &&
               std::fabs(camera.K_(1, 2) - (8.0 * 9 / 17 - 0.5)) < 1e-5,
        "colmap: principal point scaled and shifted by -0.5");
    const Image8& image = scene.train_images_[0];
    Expect(image.width == 10 && image.height == 9 && image.data[0] == 30 && image.data[1] == 20 && image.data[2] == 10,
        "colmap: pixels in RGB order");
    Expect(camera.R_c_w_(0, 0) == 1.0f && camera.R_c_w_(1, 1) == 1.0f && camera.R_c_w_(0, 1) == 0.0f,
        "colmap: identity quaternion gives the identity rotation");
    Expect(scene.point_cloud_.positions_.size() == 4 && scene.point_cloud_.colors_.size() == 4, "colmap: 4 points");
    Expect(std::fabs(scene.point_cloud_.colors_[0][0] - 100.0f / 255.0f) < 1e-6 &&
               std::fabs(scene.point_cloud_.positions_[3][0] - 0.1f) < 1e-7 &&
               std::fabs(scene.point_cloud_.positions_[3][2] - 3.0f) < 1e-7,
        "colmap: point positions and colours");
    Scene missing;
    Expect(!ReadColmapScene(root + "/does-not-exist", 1, missing), "colmap: a missing scene is rejected");
    Scene truncated;
    Expect(!ReadColmapScene(root + "/colmap_truncated", 1, truncated), "colmap: a truncated images.bin is rejected");
}

void Codecs(const std::string& root) {
    for (const char* name : {"rgb.png", "rgba.png", "grey.png", "filtered.png", "plain.ppm"}) {
        Image8 image;
        bool read = ReadImageRgb(root + "/codec/" + name, image);
        std::vector<uint8_t> expected = ReadBytes(root + "/codec/" + std::string(name) + ".rgb");
        std::string what = std::string("codec: ") + name + " decodes to the expected pixels";
        Expect(read && image.data == expected, what.c_str());
    }
    Image8 jpeg;
    Expect(!ReadImageRgb(root + "/codec/not_supported.jpg", jpeg), "codec: JPEG is rejected");
    Image8 source;
    Expect(ReadImageRgb(root + "/codec/area_source.png", source), "area: source decodes");
    std::vector<uint8_t> expected = ReadBytes(root + "/codec/area_expected.rgb");
    Image8 shrunk = DownscaleArea(source, 20, 16);
    int worst = 0;
    for (size_t i = 0; i < shrunk.data.size() && i < expected.size(); ++i) {
        worst = std::max(worst, std::abs(int(shrunk.data[i]) - int(expected[i])));
    }
    Expect(shrunk.data.size() == expected.size() && worst <= 1, "area: 61x47 -> 20x16 within 1 of the exact area mean");
    // Integer factor: exact means of 2x2 blocks.
    Image8 blocks;
    blocks.width = 4;
    blocks.height = 2;
    blocks.data = {0, 0, 0, 10, 20, 30, 100, 100, 100, 7, 7, 7, 2, 4, 6, 12, 22, 33, 1, 1, 1, 2, 2, 2};
    Image8 half = DownscaleArea(blocks, 2, 1);
    Expect(half.data == std::vector<uint8_t>({6, 12, 17, 28, 28, 28}), "area: 2x2 block means, halves to even");
    Image8 round_trip;
    Expect(WritePngRgb(root + "/written.png", source) && ReadImageRgb(root + "/written.png", round_trip) &&
               round_trip.data == source.data,
        "codec: PNG written here reads back unchanged");
}

void Ply(const std::string& root) {
    std::vector<Gaussian> expected(3);
    for (int i = 0; i < 3; ++i) {
        Gaussian& g = expected[static_cast<size_t>(i)];
        g.position_ = MakeVector3f(0.1f * i, -0.2f * i, 3.0f + i);
        g.log_scale_ = MakeVector3f(-2.0f, -2.5f + 0.1f * i, -3.0f);
        g.opacity_logit_ = 0.3f * i - 0.5f;
        for (int k = 0; k < 16; ++k) {
            for (int c = 0; c < 3; ++c) {
                g.sh_(k, c) = 0.01f * (k + 1) * (c + 1) * (i + 1);
            }
        }
    }
    expected[1].rotation_ = Vector3f::Zero();
    expected[2].rotation_ = MakeVector3f(1.0f, -0.7f, 0.2f);
    expected[0].rotation_ = MakeVector3f(0.15f, -0.2f, 0.1f);
    const std::string path = root + "/model.ply";
    std::vector<Gaussian> actual;
    Expect(WriteGaussiansPly(path, expected) && ReadGaussiansPly(path, actual) && actual.size() == 3,
        "ply: three Gaussians round trip");
    bool same = actual.size() == 3;
    for (size_t i = 0; same && i < 3; ++i) {
        for (int a = 0; a < 3; ++a) {
            same = same && actual[i].position_[a] == expected[i].position_[a] &&
                   actual[i].log_scale_[a] == expected[i].log_scale_[a] &&
                   std::fabs(actual[i].rotation_[a] - expected[i].rotation_[a]) < 2e-6f;
        }
        same = same && actual[i].opacity_logit_ == expected[i].opacity_logit_;
        for (int k = 0; k < 16; ++k) {
            for (int c = 0; c < 3; ++c) {
                same = same && actual[i].sh_(k, c) == expected[i].sh_(k, c);
            }
        }
    }
    Expect(same, "ply: parameters preserved, rotations through the quaternion to 2e-6");
    std::filesystem::resize_file(path, std::filesystem::file_size(path) - 4);
    std::vector<Gaussian> truncated;
    Expect(!ReadGaussiansPly(path, truncated) && truncated.empty(), "ply: a truncated file is rejected");
    Expect(!ReadGaussiansPly(root + "/missing.ply", truncated), "ply: a missing file is rejected");
}

} // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: io_test <fixture directory>\n");
        return 2;
    }
    const std::string root = argv[1];
    ColmapFixture(root);
    Codecs(root);
    Ply(root);
    std::printf("%d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
