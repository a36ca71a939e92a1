// Renderer (reference src/core/render/renderer.h:10-89): projection + SH colour, tile lists and
// the tile rasteriser, over the C ABI in include/gsn_hip.h. Getters return device pointers that
// stay valid until the next Render. GetTileRects and GetRecordCount are additions used by the
// backward pass.
#pragma once

#include "core/types/buffers.h"

#include <hip/hip_vector_types.h>

#include <cstddef>
#include <cstdint>

namespace gs_native {

// The per-Gaussian outputs of a projection (device pointers; see Renderer::ProjectionBuffers).
struct ProjectionTargets {
    float* means = nullptr;
    float* conic_opacities = nullptr;
    float* colors = nullptr;
    float* depths = nullptr;
    int* tile_rects = nullptr;
    int* record_offsets = nullptr;
};

class Renderer {
public:
    Renderer() = default;
    ~Renderer();

    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    bool Render(
        const GaussianBuffers& gaussians,
        const CameraParams& camera,
        const float* background,
        float* image,
        int sh_degree);

    // Additions for the fused training step: the per-Gaussian projection buffers for `count`
    // Gaussians (grown if needed) for a producer other than Render (Solver::IterateAndProject),
    // and Render from a projection already written there (tile lists and rasteriser only).
    bool ProjectionBuffers(int count, ProjectionTargets* targets);
    bool RenderProjected(
        const GaussianBuffers& gaussians,
        const CameraParams& camera,
        const float* background,
        float* image);

    const float2* GetMeans() const {
        return means_;
    }

    const float4* GetConicOpacities() const {
        return conic_opacities_;
    }

    const float3* GetColors() const {
        return colors_;
    }

    const int* GetTileGaussianIds() const {
        return tile_lists_.gaussian_ids_;
    }

    const int64_t* GetTileGaussianOffsets() const {
        return tile_lists_.gaussian_offsets_;
    }

    const int* GetRecordOffsets() const {
        return record_offsets_;
    }

    const int4* GetTileRects() const {
        return tile_rects_;
    }

    // Camera-space depth per Gaussian of the last Render (addition: the learned densification
    // features read it; the reference keeps depths internal).
    const float* GetDepths() const {
        return depths_;
    }

    int GetRecordCount() const {
        return record_count_;
    }

    int GetGaussianCapacity() const {
        return gaussian_capacity_;
    }

    int GetTileCapacity() const {
        return tile_capacity_;
    }

private:
    bool ProjectGaussians(const GaussianBuffers& gaussians, const CameraParams& camera, int sh_degree);
    bool EnsureGaussianCapacity(int count);
    bool BuildTileLists(const GaussianBuffers& gaussians, const CameraParams& camera);
    bool RenderTiles(const CameraParams& camera, const float* background, float* image);
    void ReleaseGaussianBuffers();
    void ReleaseRecordBuffers();

    float2* means_ = nullptr;
    float4* conic_opacities_ = nullptr;
    float3* colors_ = nullptr;
    float* depths_ = nullptr;
    int4* tile_rects_ = nullptr;
    int* record_offsets_ = nullptr;
    int gaussian_capacity_ = 0;

    TileListBuffers tile_lists_;
    int tile_capacity_ = 0;
    int record_capacity_ = 0;
    int record_count_ = 0;
};

} // namespace gs_native
