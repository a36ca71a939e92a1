// Tile rasteriser forward (reference src/core/render/renderer.cu:273-338).
// Per pixel, front to back: skip power > 0; alpha = min(0.99, o * e^power); skip alpha < 1/255;
// if T * (1 - alpha) < 1e-4 the pixel is done without blending that splat; otherwise
// C += c * alpha * T and T *= 1 - alpha. Output = C + T * background.
#include "raster_common.h"

namespace gsn {
namespace {

// Store = true additionally writes, per pixel, the final transmittance and the tile-relative
// index + 1 of the last splat blended (0 when none): the state a back-to-front backward needs.
// Count = true adds the number of (pixel, splat) pairs evaluated (pixel not yet done, splat's
// band kept by the mask) to *pair_count: the operation count of the VALU roofline.
template<bool Store, bool Count>
__global__ __launch_bounds__(64) void raster_forward_kernel(const int* __restrict__ ids,
    const int64_t* __restrict__ tile_offsets, const float2* __restrict__ means, const float4* __restrict__ conics,
    const float* __restrict__ colors, Camera cam, float* __restrict__ image, float bg_r, float bg_g, float bg_b,
    int tile_count, float* __restrict__ final_transmittance, int* __restrict__ last_contributor,
    unsigned long long* __restrict__ pair_count) {
    const int tile = xcd_remap(blockIdx.x, tile_count);
    const int lane = threadIdx.x;
    const int x0 = (tile % cam.tile_count_x) * kTileSize;
    const int y0 = (tile / cam.tile_count_x) * kTileSize;
    const int px = x0 + (lane & 15);
    const int py = y0 + (lane >> 4);
    const float fx = static_cast<float>(px);
    float fy[4], T[4], cr[4], cg[4], cb[4];
    int last[4] = {0, 0, 0, 0};
    unsigned evaluated = 0;
    bool done[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int y = py + 4 * b;
        fy[b] = static_cast<float>(y);
        done[b] = !(px < cam.width && y < cam.height);
        T[b] = 1.0f;
        cr[b] = 0.0f;
        cg[b] = 0.0f;
        cb[b] = 0.0f;
    }
    const int64_t begin = tile_offsets[tile];
    const int64_t end = tile_offsets[tile + 1];
    for (int64_t base = begin; base < end; base += 64) {
        int active = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (__ballot(!done[b])) {
                active |= 1 << b;
            }
        }
        if (active == 0) {
            break;
        }
        const int64_t record = base + lane;
        float mx = 0.0f, my = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        float4 conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int mask = 0;
        if (record < end) {
            int id = ids[record];
            float2 mean = means[id];
            conic = conics[id];
            const float* color = colors + static_cast<int64_t>(id) * 3;
            sr = color[0];
            sg = color[1];
            sb = color[2];
            mx = mean.x;
            my = mean.y;
            mask = band_mask(mx, my, conic, x0, y0);
        }
        uint64_t pending = __ballot((mask & active) != 0);
        while (pending) {
            const int j = __builtin_ctzll(pending);
            pending &= pending - 1;
            const int smask = readlane_i(mask, j) & active;
            const float sx = readlane_f(mx, j);
            const float sy = readlane_f(my, j);
            const float ca = readlane_f(conic.x, j);
            const float cbv = readlane_f(conic.y, j);
            const float cc = readlane_f(conic.z, j);
            const float op = readlane_f(conic.w, j);
            const float r = readlane_f(sr, j);
            const float g = readlane_f(sg, j);
            const float bl = readlane_f(sb, j);
            const float dx = sx - fx;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (smask & (1 << b)) {
                    if (Count) {
                        evaluated += done[b] ? 0u : 1u;
                    }
                    float dy = sy - fy[b];
                    float power = -0.5f * (ca * dx * dx + cc * dy * dy) - cbv * dx * dy;
                    float alpha = fminf(kAlphaMax, op * __expf(power));
                    bool take = !done[b] && !(power > 0.0f) && alpha >= kAlphaMin;
                    float next = T[b] * (1.0f - alpha);
                    bool finish = take && next < kTransmittanceMin;
                    bool blend = take && !finish;
                    float w = alpha * T[b];
                    if (blend) {
                        cr[b] += r * w;
                        cg[b] += g * w;
                        cb[b] += bl * w;
                        T[b] = next;
                        if (Store) {
                            last[b] = static_cast<int>(base - begin) + j + 1;
                        }
                    }
                    done[b] = done[b] || finish;
                }
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int y = py + 4 * b;
        if (px < cam.width && y < cam.height) {
            int64_t index = static_cast<int64_t>(y) * cam.width + px;
            float* pixel = image + index * 3;
            pixel[0] = cr[b] + T[b] * bg_r;
            pixel[1] = cg[b] + T[b] * bg_g;
            pixel[2] = cb[b] + T[b] * bg_b;
            if (Store) {
                final_transmittance[index] = T[b];
                last_contributor[index] = last[b];
            }
        }
    }
    if (Count) {
        unsigned long long total = evaluated;
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1) {
            total += __shfl_xor(total, offset);
        }
        if (lane == 0) {
            atomicAdd(pair_count, total);
        }
    }
}

} // namespace
} // namespace gsn

namespace gsn {
namespace {

int launch_forward(gsn_context* context, const Camera& cam, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, const float* means, const float* conic_opacities, const float* colors,
    const float* background, float* image, float* final_transmittance, int* last_contributor, unsigned tiles);

int render_tiles(gsn_context* context, const gsn_camera* camera, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, const float* means, const float* conic_opacities, const float* colors,
    const float* background, float* image, float* final_transmittance, int* last_contributor) {
    if (!context || !camera || !tile_offsets || !background || !image || camera->width <= 0 || camera->height <= 0 ||
        camera->tile_count_x <= 0 || camera->tile_count_y <= 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_render_tiles: invalid arguments");
    }
    int64_t tiles = static_cast<int64_t>(camera->tile_count_x) * camera->tile_count_y;
    if (tiles >= 0x7fffffff) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_render_tiles: too many tiles");
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    bool forked = false;
    if (context->prefetch_armed) {
        // Armed by gsn_context_prefetch_backward: the backward's preparation of this projection
        // runs on the auxiliary stream beside this kernel and the image gradient. The fork point
        // is recorded before the kernel and the auxiliary launches are issued after it, so the
        // host's work for them does not delay the forward.
        context->prefetch_armed = false;
        if (means == context->projected_means && conic_opacities == context->projected_conics &&
            colors == context->projected_colors) {
            GSN_TRY(fork_backward_prefetch(context, cam, &forked));
        }
    }
    GSN_TRY(launch_forward(context, cam, tile_gaussian_ids, tile_offsets, means, conic_opacities, colors, background,
        image, final_transmittance, last_contributor, static_cast<unsigned>(tiles)));
    return forked ? queue_backward_prefetch(context, cam, tile_offsets) : GSN_OK;
}

int launch_forward(gsn_context* context, const Camera& cam, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, const float* means, const float* conic_opacities, const float* colors,
    const float* background, float* image, float* final_transmittance, int* last_contributor, unsigned tiles) {
    unsigned long long* pair_count = context->count_raster_pairs ? context->raster_pairs : nullptr;
    StageScope scope(context, kStageRasterForward);
    const unsigned grid = static_cast<unsigned>(tiles);
    const float2* m = reinterpret_cast<const float2*>(means);
    const float4* q = reinterpret_cast<const float4*>(conic_opacities);
    if (final_transmittance && pair_count) {
        raster_forward_kernel<true, true><<<grid, 64, 0, context->stream>>>(tile_gaussian_ids, tile_offsets, m, q,
            colors, cam, image, background[0], background[1], background[2], static_cast<int>(tiles),
            final_transmittance, last_contributor, pair_count);
    }
    else if (final_transmittance) {
        raster_forward_kernel<true, false><<<grid, 64, 0, context->stream>>>(tile_gaussian_ids, tile_offsets, m, q,
            colors, cam, image, background[0], background[1], background[2], static_cast<int>(tiles),
            final_transmittance, last_contributor, nullptr);
    }
    else if (pair_count) {
        raster_forward_kernel<false, true><<<grid, 64, 0, context->stream>>>(tile_gaussian_ids, tile_offsets, m, q,
            colors, cam, image, background[0], background[1], background[2], static_cast<int>(tiles), nullptr,
            nullptr, pair_count);
    }
    else {
        raster_forward_kernel<false, false><<<grid, 64, 0, context->stream>>>(tile_gaussian_ids, tile_offsets, m, q,
            colors, cam, image, background[0], background[1], background[2], static_cast<int>(tiles), nullptr,
            nullptr, nullptr);
    }
    return check_launch("raster_forward_kernel");
}

} // namespace
} // namespace gsn

extern "C" int gsn_render_tiles(gsn_context* context, const gsn_camera* camera, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, const float* means, const float* conic_opacities, const float* colors,
    const float* background, float* image) {
    return gsn::render_tiles(context, camera, tile_gaussian_ids, tile_offsets, means, conic_opacities, colors,
        background, image, nullptr, nullptr);
}

extern "C" int gsn_render_tiles_stored(gsn_context* context, const gsn_camera* camera, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, const float* means, const float* conic_opacities, const float* colors,
    const float* background, float* image, float* final_transmittance, int32_t* last_contributor) {
    if (!final_transmittance || !last_contributor) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_render_tiles_stored: null per-pixel state");
    }
    return gsn::render_tiles(context, camera, tile_gaussian_ids, tile_offsets, means, conic_opacities, colors,
        background, image, final_transmittance, last_contributor);
}
