// Renderer over the C ABI (reference src/core/render/renderer.cu:342-590).
#include "core/render/renderer.h"

#include "device_memory.h"
#include "util/device.h"

#include <algorithm>
#include <climits>

namespace gs_native {

Renderer::~Renderer() {
    ReleaseGaussianBuffers();
    ReleaseRecordBuffers();
    DeviceFree(tile_lists_.gaussian_offsets_);
}

void Renderer::ReleaseGaussianBuffers() {
    DeviceFree(means_);
    DeviceFree(conic_opacities_);
    DeviceFree(colors_);
    DeviceFree(depths_);
    DeviceFree(tile_rects_);
    DeviceFree(record_offsets_);
    means_ = nullptr;
    conic_opacities_ = nullptr;
    colors_ = nullptr;
    depths_ = nullptr;
    tile_rects_ = nullptr;
    record_offsets_ = nullptr;
    gaussian_capacity_ = 0;
}

void Renderer::ReleaseRecordBuffers() {
    DeviceFree(tile_lists_.gaussian_ids_);
    tile_lists_.gaussian_ids_ = nullptr;
    tile_lists_.capacity_ = 0;
    record_capacity_ = 0;
}

bool Renderer::Render(
    const GaussianBuffers& gaussians,
    const CameraParams& camera,
    const float* background,
    float* image,
    int sh_degree) {
    if (!ProjectGaussians(gaussians, camera, sh_degree)) {
        return false;
    }
    if (!BuildTileLists(gaussians, camera)) {
        return false;
    }
    return RenderTiles(camera, background, image);
}

bool Renderer::RenderProjected(
    const GaussianBuffers& gaussians,
    const CameraParams& camera,
    const float* background,
    float* image) {
    return BuildTileLists(gaussians, camera) && RenderTiles(camera, background, image);
}

bool Renderer::ProjectionBuffers(int count, ProjectionTargets* targets) {
    if (!targets || count < 0 || !EnsureGaussianCapacity(count)) {
        return false;
    }
    targets->means = reinterpret_cast<float*>(means_);
    targets->conic_opacities = reinterpret_cast<float*>(conic_opacities_);
    targets->colors = reinterpret_cast<float*>(colors_);
    targets->depths = depths_;
    targets->tile_rects = reinterpret_cast<int*>(tile_rects_);
    targets->record_offsets = record_offsets_;
    return true;
}

bool Renderer::EnsureGaussianCapacity(int count) {
    if (count > gaussian_capacity_ || !record_offsets_) {
        ReleaseGaussianBuffers();
        size_t n = static_cast<size_t>(count);
        bool ok = DeviceAllocateCount(&record_offsets_, n + 1);
        if (ok && count > 0) {
            ok = DeviceAllocateCount(&means_, n) && DeviceAllocateCount(&conic_opacities_, n) &&
                DeviceAllocateCount(&colors_, n) && DeviceAllocateCount(&depths_, n) &&
                DeviceAllocateCount(&tile_rects_, n);
        }
        if (!ok) {
            ReleaseGaussianBuffers();
            return false;
        }
        gaussian_capacity_ = count;
    }
    return true;
}

bool Renderer::ProjectGaussians(const GaussianBuffers& gaussians, const CameraParams& camera, int sh_degree) {
    int count = gaussians.gaussian_count_;
    gsn_context* context = DefaultContext();
    if (!context || count < 0 || (count > 0 && !gaussians.gaussians_) || !EnsureGaussianCapacity(count)) {
        return false;
    }
    sh_degree = std::clamp(sh_degree, 0, 3);
    return Check(gsn_project_gaussians(context, gaussians.gaussians_, count, AsCamera(&camera), sh_degree,
                     reinterpret_cast<float*>(means_), reinterpret_cast<float*>(conic_opacities_),
                     reinterpret_cast<float*>(colors_), depths_, reinterpret_cast<int*>(tile_rects_), record_offsets_),
        "gsn_project_gaussians");
}

bool Renderer::BuildTileLists(const GaussianBuffers& gaussians, const CameraParams& camera) {
    int count = gaussians.gaussian_count_;
    int64_t tile_count_long = static_cast<int64_t>(camera.tile_count_x_) * camera.tile_count_y_;
    gsn_context* context = DefaultContext();
    if (!context || count < 0 || count == INT_MAX || count > gaussian_capacity_ || !record_offsets_ ||
        camera.tile_count_x_ <= 0 || camera.tile_count_y_ <= 0 || tile_count_long >= INT_MAX) {
        return false;
    }
    int tile_count = static_cast<int>(tile_count_long);
    if (tile_count > tile_capacity_) {
        DeviceFree(tile_lists_.gaussian_offsets_);
        tile_lists_.gaussian_offsets_ = nullptr;
        tile_capacity_ = 0;
        if (!DeviceAllocateCount(&tile_lists_.gaussian_offsets_, static_cast<size_t>(tile_count) + 1)) {
            return false;
        }
        tile_capacity_ = tile_count;
    }
    int records = 0;
    if (!Check(gsn_scan_record_offsets(context, record_offsets_, count, &records), "gsn_scan_record_offsets")) {
        return false;
    }
    if (records > record_capacity_) {
        ReleaseRecordBuffers();
        if (!DeviceAllocateCount(&tile_lists_.gaussian_ids_, static_cast<size_t>(records))) {
            return false;
        }
        record_capacity_ = records;
        tile_lists_.capacity_ = records;
    }
    record_count_ = records;
    return Check(gsn_build_tile_lists(context, AsCamera(&camera), count, reinterpret_cast<const int*>(tile_rects_),
                     depths_, record_offsets_, records, tile_lists_.gaussian_ids_, tile_lists_.gaussian_offsets_,
                     nullptr),
        "gsn_build_tile_lists");
}

bool Renderer::RenderTiles(const CameraParams& camera, const float* background, float* image) {
    int64_t tile_count = static_cast<int64_t>(camera.tile_count_x_) * camera.tile_count_y_;
    gsn_context* context = DefaultContext();
    if (!context || !background || !image || !tile_lists_.gaussian_offsets_ || camera.width_ <= 0 ||
        camera.height_ <= 0 || camera.tile_count_x_ <= 0 || camera.tile_count_y_ <= 0 || tile_count > tile_capacity_) {
        return false;
    }
    return Check(gsn_render_tiles(context, AsCamera(&camera), tile_lists_.gaussian_ids_, tile_lists_.gaussian_offsets_,
                     reinterpret_cast<const float*>(means_), reinterpret_cast<const float*>(conic_opacities_),
                     reinterpret_cast<const float*>(colors_), background, image),
        "gsn_render_tiles");
}

} // namespace gs_native
