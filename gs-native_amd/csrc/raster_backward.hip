// Tile rasteriser backward (reference src/core/solver/solver.cu:325-466), recompute form.
//
// Nothing is stored per pixel by the forward pass: each pixel replays its splats front to back
// from T = 1 and uses the final colour C (image) to recover the colour behind every splat:
//   dL/dalpha_k = sum_c g_c (prefix_c + T_k c_c - C_c) / (1 - alpha_k),
// with prefix_c = sum_{i<k} c_i alpha_i T_i. Geometry and opacity gradients are taken only when
// the raw alpha is below the 0.99 cap (solver.cu:398-408).
//
// Gradient accumulation without global atomics: the 64 lanes' contributions for a splat are
// summed with wave_sum12 and the tile's total for that (tile, Gaussian) record is written to a
// per-record partial slot. The slot index is the record's position in the Gaussian-id-ordered
// enumeration, record_offsets[g] + row * width + column of the tile inside g's rectangle, so
// each Gaussian's partials are contiguous and a later per-Gaussian pass reduces them in a fixed
// order (deterministic results).
#include "raster_common.h"

namespace gsn {

namespace {

__device__ __forceinline__ int64_t partial_slot(int id, int tile_x, int tile_y, const int4* rects,
    const int* record_offsets) {
    int4 rect = rects[id];
    int width = rect.z - rect.x;
    return static_cast<int64_t>(record_offsets[id]) + (tile_y - rect.y) * width + (tile_x - rect.x);
}

__global__ __launch_bounds__(64) void raster_backward_kernel(const int* __restrict__ ids,
    const int64_t* __restrict__ tile_offsets, const float2* __restrict__ means, const float4* __restrict__ conics,
    const float* __restrict__ colors, const int4* __restrict__ rects, const int* __restrict__ record_offsets,
    Camera cam, const float* __restrict__ image, const float* __restrict__ image_gradients, float* __restrict__ partials,
    int tile_count, int record_count) {
    __shared__ float s_partial[64 * kPartialFloats];
    const int tile = xcd_remap(blockIdx.x, tile_count);
    const int lane = threadIdx.x;
    const int tile_x = tile % cam.tile_count_x;
    const int tile_y = tile / cam.tile_count_x;
    const int x0 = tile_x * kTileSize;
    const int y0 = tile_y * kTileSize;
    const int px = x0 + (lane & 15);
    const int py = y0 + (lane >> 4);
    const float fx = static_cast<float>(px);
    float fy[4], T[4], final_color[4][3], grad[4][3], prefix[4][3];
    bool done[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int y = py + 4 * b;
        bool inside = px < cam.width && y < cam.height;
        fy[b] = static_cast<float>(y);
        done[b] = !inside;
        T[b] = 1.0f;
        int64_t pixel = (static_cast<int64_t>(y) * cam.width + px) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            final_color[b][c] = inside ? image[pixel + c] : 0.0f;
            grad[b][c] = inside ? image_gradients[pixel + c] : 0.0f;
            prefix[b][c] = 0.0f;
        }
    }
    const int64_t begin = tile_offsets[tile];
    const int64_t end = tile_offsets[tile + 1];
    int64_t base = begin;
    for (; base < end; base += 64) {
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
        const bool valid = record < end;
        float mx = 0.0f, my = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        float4 conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int mask = 0;
        int64_t slot = 0;
        if (valid) {
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
            slot = partial_slot(id, tile_x, tile_y, rects, record_offsets);
        }
#pragma unroll
        for (int k = 0; k < kPartialFloats; ++k) {
            s_partial[lane * kPartialFloats + k] = 0.0f;
        }
        __syncthreads();
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
            const float col[3] = {readlane_f(sr, j), readlane_f(sg, j), readlane_f(sb, j)};
            const float dx = sx - fx;
            float acc[12];
#pragma unroll
            for (int k = 0; k < 12; ++k) {
                acc[k] = 0.0f;
            }
            bool contributed = false;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (smask & (1 << b)) {
                    float dy = sy - fy[b];
                    float power = -0.5f * (ca * dx * dx + cc * dy * dy) - cbv * dx * dy;
                    float raw_alpha = op * __expf(power);
                    float alpha = fminf(kAlphaMax, raw_alpha);
                    bool take = !done[b] && !(power > 0.0f) && alpha >= kAlphaMin;
                    float next = T[b] * (1.0f - alpha);
                    bool finish = take && next < kTransmittanceMin;
                    bool blend = take && !finish;
                    done[b] = done[b] || finish;
                    if (blend) {
                        contributed = true;
                        float weight = alpha * T[b];
                        float inv = __builtin_amdgcn_rcpf(1.0f - alpha);
                        float alpha_gradient = 0.0f;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            alpha_gradient += grad[b][c] * (prefix[b][c] + T[b] * col[c] - final_color[b][c]) * inv;
                            acc[6 + c] += weight * grad[b][c];
                        }
                        if (raw_alpha < kAlphaMax) {
                            float common = alpha_gradient * raw_alpha;
                            float qx = ca * dx + cbv * dy;
                            float qy = cbv * dx + cc * dy;
                            float gx = -common * qx;
                            float gy = -common * qy;
                            acc[0] += gx;
                            acc[1] += gy;
                            acc[2] += 0.5f * common * qx * qx;
                            acc[3] += common * qx * qy;
                            acc[4] += 0.5f * common * qy * qy;
                            acc[5] += common * (1.0f - op);
                            acc[9] += fabsf(gx);
                            acc[10] += fabsf(gy);
                        }
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            prefix[b][c] += col[c] * weight;
                        }
                        T[b] = next;
                    }
                }
            }
            if (__ballot(contributed)) {
                float sums[3];
                wave_sum12(acc, sums);
                if ((lane & 15) == 0) {
                    int row = lane >> 4;
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        s_partial[j * kPartialFloats + 3 * row + k] = sums[k];
                    }
                }
            }
        }
        __syncthreads();
        if (valid && slot >= 0 && slot < record_count) {
            const float4* source = reinterpret_cast<const float4*>(s_partial + lane * kPartialFloats);
            float4* destination = reinterpret_cast<float4*>(partials + slot * kPartialFloats);
            destination[0] = source[0];
            destination[1] = source[1];
            destination[2] = source[2];
        }
        __syncthreads();
    }
    // Records after every pixel of the tile terminated contribute nothing.
    for (; base < end; base += 64) {
        const int64_t record = base + lane;
        if (record < end) {
            int id = ids[record];
            int64_t slot = partial_slot(id, tile_x, tile_y, rects, record_offsets);
            if (slot < 0 || slot >= record_count) {
                continue;
            }
            float4* destination = reinterpret_cast<float4*>(partials + slot * kPartialFloats);
            float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            destination[0] = zero;
            destination[1] = zero;
            destination[2] = zero;
        }
    }
}

__global__ __launch_bounds__(64) void wave_sum12_selftest_kernel(const float* input, float* output) {
    int lane = threadIdx.x;
    float v[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        v[k] = input[lane * 12 + k];
    }
    float sums[3];
    wave_sum12(v, sums);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        output[lane * 3 + k] = sums[k];
    }
}

} // namespace

int launch_splat_backward(gsn_context* context, const Camera& cam, const int* ids, const int64_t* tile_offsets,
    const float* means, const float* conics, const float* colors, const int* rects, const int* record_offsets,
    const float* image, const float* image_gradients, float* partials, int record_count) {
    int64_t tiles = static_cast<int64_t>(cam.tile_count_x) * cam.tile_count_y;
    StageScope scope(context, kStageRasterBackward);
    raster_backward_kernel<<<static_cast<unsigned>(tiles), 64, 0, context->stream>>>(ids, tile_offsets,
        reinterpret_cast<const float2*>(means), reinterpret_cast<const float4*>(conics), colors,
        reinterpret_cast<const int4*>(rects), record_offsets, cam, image, image_gradients, partials,
        static_cast<int>(tiles), record_count);
    return check_launch("raster_backward_kernel");
}

} // namespace gsn

// Diagnostic: runs the wave reduction on one wave. input float[64][12] -> output float[64][3].
extern "C" int gsn_selftest_wave_sum12(gsn_context* context, const float* input, float* output) {
    using namespace gsn;
    if (!context || !input || !output) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_selftest_wave_sum12: null argument");
    }
    wave_sum12_selftest_kernel<<<1, 64, 0, context->stream>>>(input, output);
    return check_launch("wave_sum12_selftest_kernel");
}
