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

// Everything the backward needs of one Gaussian in 48 bytes, so that each record costs one
// gather instead of five (means, conics, colours, rectangles, record offsets):
//   a = (mean x, mean y, conic A, conic B), b = (conic C, opacity, red, green),
//   c = (blue, record offset, rectangle x0 | y0 << 16, rectangle width) (last three as int bits).
struct alignas(16) PackedSplat {
    float4 a, b, c;
};

__global__ __launch_bounds__(256) void pack_splats_kernel(const float2* __restrict__ means,
    const float4* __restrict__ conics, const float* __restrict__ colors, const int4* __restrict__ rects,
    const int* __restrict__ record_offsets, int count, PackedSplat* __restrict__ packed) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count || record_offsets[id + 1] == record_offsets[id]) {
        return;  // culled: never referenced by a record
    }
    float2 mean = means[id];
    float4 conic = conics[id];
    int4 rect = rects[id];
    const float* color = colors + static_cast<int64_t>(id) * 3;
    PackedSplat out;
    out.a = make_float4(mean.x, mean.y, conic.x, conic.y);
    out.b = make_float4(conic.z, conic.w, color[0], color[1]);
    out.c = make_float4(color[2], __int_as_float(record_offsets[id]), __int_as_float(rect.x | (rect.y << 16)),
        __int_as_float(rect.z - rect.x));
    packed[id] = out;
}

__device__ __forceinline__ int64_t partial_slot(float4 c, int tile_x, int tile_y) {
    int origin = __float_as_int(c.z);
    int width = __float_as_int(c.w);
    return static_cast<int64_t>(__float_as_int(c.y)) + (tile_y - (origin >> 16)) * width + (tile_x - (origin & 0xffff));
}

// Heavy-first tile schedule. The backward holds 4 waves per SIMD, so a 1080p frame's 8 160 tiles
// run in two rounds and a long tile (several times the mean record count) that starts late sets
// the kernel's end. Tiles are dispatched in descending order of record count instead (longest
// processing time first), bucketed at 4 buckets per octave. The order within a bucket depends on
// LDS atomics and is not fixed, which does not matter: each tile writes only its own pixels and
// record slots, so results do not depend on the schedule.
constexpr int kScheduleBins = 128;
constexpr int kScheduleThreads = 1024;

__device__ __forceinline__ int schedule_bin(int64_t records) {
    uint32_t v = static_cast<uint32_t>(min(max(records, int64_t{0}), int64_t{0x7ffffffe})) + 1u;
    int octave = 31 - __clz(v);
    int fraction = octave >= 2 ? static_cast<int>((v >> (octave - 2)) & 3u) : static_cast<int>((v << (2 - octave)) & 3u);
    return kScheduleBins - 1 - (octave * 4 + fraction);
}

// Heavy-first order of `units` work units; unit u owns records [offsets[u * stride],
// offsets[min((u + 1) * stride, count)]). Run by one whole workgroup. The bins of the first
// kScheduleItems * kScheduleThreads units are computed once into registers, with all their offset
// loads in flight together; units past that recompute their bin. (Matching equal bins across the
// wave to issue one LDS atomic per bin measured slower: the strided offset loads dominate.)
constexpr int kScheduleItems = 16;

template<typename Offset>
__device__ void schedule_units(const Offset* __restrict__ offsets, int stride, int count, int units,
    int* __restrict__ order, int* histogram, int* cursor) {
    auto bin = [&](int u) {
        int64_t begin = offsets[min(static_cast<int64_t>(u) * stride, static_cast<int64_t>(count))];
        int64_t end = offsets[min(static_cast<int64_t>(u + 1) * stride, static_cast<int64_t>(count))];
        return schedule_bin(end - begin);
    };
    int bins[kScheduleItems];
#pragma unroll
    for (int i = 0; i < kScheduleItems; ++i) {
        const int u = threadIdx.x + i * kScheduleThreads;
        bins[i] = u < units ? bin(u) : -1;
    }
    for (int b = threadIdx.x; b < kScheduleBins; b += kScheduleThreads) {
        histogram[b] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScheduleItems; ++i) {
        if (bins[i] >= 0) {
            atomicAdd(&histogram[bins[i]], 1);
        }
    }
    for (int u = threadIdx.x + kScheduleItems * kScheduleThreads; u < units; u += kScheduleThreads) {
        atomicAdd(&histogram[bin(u)], 1);
    }
    __syncthreads();
    // Exclusive scan of the bins by the first wave, two bins per lane.
    static_assert(kScheduleBins == 128, "the scan below covers 2 x 64 bins");
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int first = histogram[2 * lane];
        const int pair = first + histogram[2 * lane + 1];
        int inclusive = pair;
#pragma unroll
        for (int offset = 1; offset < 64; offset <<= 1) {
            const int below = __shfl_up(inclusive, offset);
            inclusive += lane >= offset ? below : 0;
        }
        cursor[2 * lane] = inclusive - pair;
        cursor[2 * lane + 1] = inclusive - pair + first;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kScheduleItems; ++i) {
        if (bins[i] >= 0) {
            order[atomicAdd(&cursor[bins[i]], 1)] = threadIdx.x + i * kScheduleThreads;
        }
    }
    for (int u = threadIdx.x + kScheduleItems * kScheduleThreads; u < units; u += kScheduleThreads) {
        order[atomicAdd(&cursor[bin(u)], 1)] = u;
    }
}

// Two schedules, one per workgroup of the launch: the backward's tiles by record count, and the
// per-Gaussian pass's workgroups (kGradientGroup consecutive Gaussians) by the length of their
// record run. Density control appends its clones and split children, high-gradient Gaussians
// whose records are all reached, at the end of the id range; in id order their workgroups, the
// longest of the pass, were dispatched last and set its end (measured 1.3 -> 2.3 ms after one
// density step at 3M).
__global__ __launch_bounds__(kScheduleThreads) void tile_schedule_kernel(const int64_t* __restrict__ tile_offsets,
    int tile_count, int* __restrict__ order, const int* __restrict__ record_offsets, int gaussian_count,
    int* __restrict__ gaussian_order) {
    __shared__ int histogram[kScheduleBins];
    __shared__ int cursor[kScheduleBins];
    if (blockIdx.x == 0) {
        schedule_units(tile_offsets, 1, tile_count, tile_count, order, histogram, cursor);
    }
    else {
        schedule_units(record_offsets, kGradientGroup, gaussian_count,
            (gaussian_count + kGradientGroup - 1) / kGradientGroup, gaussian_order, histogram, cursor);
    }
}

// Plain stores: non-temporal stores of these scattered 16-byte pieces measured 2.3x the write
// traffic and twice the kernel time (each piece left the cache on its own).
__device__ __forceinline__ void store_partial(float* destination, const float4& value) {
    *reinterpret_cast<float4*>(destination) = value;
}

__global__ __launch_bounds__(64) void raster_backward_kernel(const int* __restrict__ ids,
    const int64_t* __restrict__ tile_offsets, const PackedSplat* __restrict__ splats,
    const int* __restrict__ tile_order, Camera cam, const float* __restrict__ image,
    const float* __restrict__ image_gradients, float* __restrict__ partials, uint8_t* __restrict__ record_flags,
    uint8_t epoch, int record_count) {
    __shared__ float s_partial[64 * kPartialFloats];
    const int tile = tile_order[blockIdx.x];
    const int lane = threadIdx.x;
    const int tile_x = tile % cam.tile_count_x;
    const int tile_y = tile / cam.tile_count_x;
    const int x0 = tile_x * kTileSize;
    const int y0 = tile_y * kTileSize;
    const int px = x0 + (lane & 15);
    const int py = y0 + (lane >> 4);
    const float fx = static_cast<float>(px);
    // A pixel that is done (outside the image, or terminated) carries T = 0: every later splat
    // then finds T (1 - alpha) below the termination threshold and blends nothing.
    float fy[4], T[4], final_color[4][3], grad[4][3], prefix[4][3];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int y = py + 4 * b;
        bool inside = px < cam.width && y < cam.height;
        fy[b] = static_cast<float>(y);
        T[b] = inside ? 1.0f : 0.0f;
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
    // Software pipeline over record batches: while batch k is processed, the splat records of
    // batch k + 1 and the ids of batch k + 2 are in flight, so the two dependent gathers of a
    // batch (id, then its 48-byte record) are hidden behind a batch of compute.
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float4 current_a = zero, current_b = zero, current_c = zero;
    if (begin + lane < end) {
        const PackedSplat* first = splats + ids[begin + lane];
        current_a = first->a;
        current_b = first->b;
        current_c = first->c;
    }
    int following_id = begin + 64 + lane < end ? ids[begin + 64 + lane] : -1;
    int64_t base = begin;
    for (; base < end; base += 64) {
        int active = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (__ballot(T[b] > 0.0f)) {
                active |= 1 << b;
            }
        }
        if (active == 0) {
            break;
        }
        const float4 splat_a = current_a, splat_b = current_b, splat_c = current_c;
        if (following_id >= 0) {
            const PackedSplat* next = splats + following_id;
            current_a = next->a;
            current_b = next->b;
            current_c = next->c;
        }
        following_id = base + 128 + lane < end ? ids[base + 128 + lane] : -1;
        const int64_t record = base + lane;
        const bool valid = record < end;
        float mx = 0.0f, my = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        float4 conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int mask = 0;
        int64_t slot = 0;
        if (valid) {
            mx = splat_a.x;
            my = splat_a.y;
            conic = make_float4(splat_a.z, splat_a.w, splat_b.x, splat_b.y);
            sr = splat_b.z;
            sg = splat_b.w;
            sb = splat_c.x;
            mask = band_mask(mx, my, conic, x0, y0);
            slot = partial_slot(splat_c, tile_x, tile_y);
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
                // Opaque to the optimiser: otherwise it folds the zero into the first band's
                // update and rebuilds twelve zeros on every path that skips a band.
                asm volatile("" : "+v"(acc[k]));
            }
            bool contributed = false;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (smask & (1 << b)) {
                    float dy = sy - fy[b];
                    float power = -0.5f * (ca * dx * dx + cc * dy * dy) - cbv * dx * dy;
                    float raw_alpha = op * __expf(power);
                    float alpha = fminf(kAlphaMax, raw_alpha);
                    bool take = !(power > 0.0f) && alpha >= kAlphaMin;
                    float next = T[b] * (1.0f - alpha);
                    bool finish = take && next < kTransmittanceMin;
                    bool blend = take && !finish;
                    // Uniform branch on "some lane blends"; inside, lanes that do not blend add
                    // exact zeros (selects, not divergent ifs), which keeps the accumulators in
                    // place instead of merging per-lane control paths.
                    if (__ballot(blend)) {
                        contributed = true;
                        float weight = blend ? alpha * T[b] : 0.0f;
                        float inv = __builtin_amdgcn_rcpf(1.0f - alpha);
                        float alpha_gradient = 0.0f;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            alpha_gradient += grad[b][c] * (prefix[b][c] + T[b] * col[c] - final_color[b][c]) * inv;
                            acc[6 + c] += weight * grad[b][c];
                        }
                        float common = blend && raw_alpha < kAlphaMax ? alpha_gradient * raw_alpha : 0.0f;
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
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            prefix[b][c] += col[c] * weight;
                        }
                        T[b] = blend ? next : T[b];
                    }
                    T[b] = finish ? 0.0f : T[b];
                }
            }
            if (contributed) {
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
            float* destination = partials + slot * kPartialFloats;
            store_partial(destination, source[0]);
            store_partial(destination + 4, source[1]);
            store_partial(destination + 8, source[2]);
            record_flags[slot] = epoch;
        }
        __syncthreads();
    }
    // Records after every pixel of the tile terminated contribute nothing; they get neither a
    // partial nor this step's flag.
}

// Stored-state variant (no counterpart in the reference; kept for the memory comparison with the
// recompute form above). The forward stored, per pixel, the final transmittance T_N and the
// tile-relative index + 1 of the last blended splat. The pass walks each tile's records back to
// front from the last contributor of any of its pixels, recovering T_k = T_{k+1} / (1 - alpha_k)
// and the colour behind splat k, B_k = T_N * background + sum_{j > k} c_j alpha_j T_j, so that
//   dL/dalpha_k = sum_c g_c (T_k c_c - B_c / (1 - alpha_k)),
// the same quantity as the recompute form's sum_c g_c (prefix_c + T_k c_c - C_c) / (1 - alpha_k).
// Partials go to the same per-record slots.
__global__ __launch_bounds__(64) void raster_backward_stored_kernel(const int* __restrict__ ids,
    const int64_t* __restrict__ tile_offsets, const PackedSplat* __restrict__ splats,
    const int* __restrict__ tile_order, Camera cam, const float* __restrict__ final_transmittance,
    const int* __restrict__ last_contributor, const float* __restrict__ image_gradients, float bg_r, float bg_g,
    float bg_b, float* __restrict__ partials, uint8_t* __restrict__ record_flags, uint8_t epoch, int record_count) {
    __shared__ float s_partial[64 * kPartialFloats];
    const int tile = tile_order[blockIdx.x];
    const int lane = threadIdx.x;
    const int tile_x = tile % cam.tile_count_x;
    const int tile_y = tile / cam.tile_count_x;
    const int x0 = tile_x * kTileSize;
    const int y0 = tile_y * kTileSize;
    const int px = x0 + (lane & 15);
    const int py = y0 + (lane >> 4);
    const float fx = static_cast<float>(px);
    const float background[3] = {bg_r, bg_g, bg_b};
    float fy[4], T[4], behind[4][3], grad[4][3];
    int last[4];
    int tile_last = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        int y = py + 4 * b;
        bool inside = px < cam.width && y < cam.height;
        fy[b] = static_cast<float>(y);
        int64_t pixel = static_cast<int64_t>(y) * cam.width + px;
        T[b] = inside ? final_transmittance[pixel] : 1.0f;
        last[b] = inside ? last_contributor[pixel] : 0;
        tile_last = max(tile_last, last[b]);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            grad[b][c] = inside ? image_gradients[pixel * 3 + c] : 0.0f;
            behind[b][c] = T[b] * background[c];
        }
    }
    // Largest last contributor over the tile (wave maximum through LDS-free lane reads).
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        tile_last = max(tile_last, __shfl_xor(tile_last, offset));
    }
    const int64_t begin = tile_offsets[tile];
    const int64_t end = tile_offsets[tile + 1];
    const int64_t stop = min(end, begin + static_cast<int64_t>(tile_last));
    for (int64_t top = stop; top > begin; top -= 64) {
        const int64_t base = max(begin, top - 64);
        const int64_t record = base + lane;
        const bool valid = record < top;
        float mx = 0.0f, my = 0.0f, sr = 0.0f, sg = 0.0f, sb = 0.0f;
        float4 conic = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int mask = 0;
        int64_t slot = 0;
        if (valid) {
            const PackedSplat splat = splats[ids[record]];
            mx = splat.a.x;
            my = splat.a.y;
            conic = make_float4(splat.a.z, splat.a.w, splat.b.x, splat.b.y);
            sr = splat.b.z;
            sg = splat.b.w;
            sb = splat.c.x;
            mask = band_mask(mx, my, conic, x0, y0);
            slot = partial_slot(splat.c, tile_x, tile_y);
        }
#pragma unroll
        for (int k = 0; k < kPartialFloats; ++k) {
            s_partial[lane * kPartialFloats + k] = 0.0f;
        }
        __syncthreads();
        uint64_t pending = __ballot(mask != 0);
        while (pending) {
            const int j = 63 - __builtin_clzll(pending);
            pending &= ~(1ull << j);
            const int position = static_cast<int>(base - begin) + j;  // tile-relative record index
            const int smask = readlane_i(mask, j);
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
                    bool blend = position < last[b] && !(power > 0.0f) && alpha >= kAlphaMin;
                    if (blend) {
                        contributed = true;
                        float inv = __builtin_amdgcn_rcpf(1.0f - alpha);
                        float before = T[b] * inv;  // T_k, transmittance in front of this splat
                        float weight = alpha * before;
                        float alpha_gradient = 0.0f;
#pragma unroll
                        for (int c = 0; c < 3; ++c) {
                            alpha_gradient += grad[b][c] * (before * col[c] - behind[b][c] * inv);
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
                            behind[b][c] += col[c] * weight;
                        }
                        T[b] = before;
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
            float* destination = partials + slot * kPartialFloats;
            store_partial(destination, source[0]);
            store_partial(destination + 4, source[1]);
            store_partial(destination + 8, source[2]);
            record_flags[slot] = epoch;
        }
        __syncthreads();
    }
    // Records behind every pixel's last contributor contribute nothing; they get neither a
    // partial nor this step's flag.
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

__global__ __launch_bounds__(256) void clear_flags_kernel(uint32_t* flags, int64_t words) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < words) {
        flags[i] = 0u;
    }
}

} // namespace

// The backward's splat copy and its two schedules (tile order, per-Gaussian group order).
struct BackwardBuffers {
    PackedSplat* splats = nullptr;
    int* order = nullptr;
    int* groups_order = nullptr;
};

static int backward_buffers(gsn_context* context, const Camera& cam, int count, BackwardBuffers* buffers) {
    const int64_t tiles = static_cast<int64_t>(cam.tile_count_x) * cam.tile_count_y;
    const int groups = (count + kGradientGroup - 1) / kGradientGroup;
    buffers->splats = static_cast<PackedSplat*>(
        scratch(context, kSlotSplats, static_cast<size_t>(count > 0 ? count : 1) * sizeof(PackedSplat)));
    buffers->order = static_cast<int*>(scratch(context, kSlotTileOrder, static_cast<size_t>(tiles) * 4));
    buffers->groups_order =
        static_cast<int*>(scratch(context, kSlotGaussianOrder, static_cast<size_t>(groups > 0 ? groups : 1) * 4));
    return buffers->splats && buffers->order && buffers->groups_order ? GSN_OK : GSN_ERROR_OUT_OF_MEMORY;
}

static int launch_backward_preparation(hipStream_t stream, const Camera& cam, int count, const int64_t* tile_offsets,
    const float* means, const float* conics, const float* colors, const int* rects, const int* record_offsets,
    const BackwardBuffers& buffers) {
    const int64_t tiles = static_cast<int64_t>(cam.tile_count_x) * cam.tile_count_y;
    pack_splats_kernel<<<grid_for(count, 256), 256, 0, stream>>>(reinterpret_cast<const float2*>(means),
        reinterpret_cast<const float4*>(conics), colors, reinterpret_cast<const int4*>(rects), record_offsets, count,
        buffers.splats);
    GSN_TRY(check_launch("pack_splats_kernel"));
    tile_schedule_kernel<<<2, kScheduleThreads, 0, stream>>>(tile_offsets, static_cast<int>(tiles), buffers.order,
        record_offsets, count, buffers.groups_order);
    return check_launch("tile_schedule_kernel");
}

int join_backward_prefetch(gsn_context* context) {
    if (!context->prefetch_pending) {
        return GSN_OK;
    }
    context->prefetch_pending = false;
    return check_hip(hipStreamWaitEvent(context->stream, context->join_event, 0), "hipStreamWaitEvent");
}

int fork_backward_prefetch(gsn_context* context, const Camera& cam, bool* forked) {
    *forked = false;
    GSN_TRY(join_backward_prefetch(context));
    const int count = context->projected_gaussians;
    if (count <= 0 || !context->aux_stream || cam.tile_count_x > 0xffff || cam.tile_count_y > 0x7fff) {
        return GSN_OK;
    }
    // Allocations first: growing a slot synchronises both streams before it frees the old buffer.
    BackwardBuffers buffers;
    GSN_TRY(backward_buffers(context, cam, count, &buffers));
    GSN_HIP(hipEventRecord(context->fork_event, context->stream));
    *forked = true;
    return GSN_OK;
}

int queue_backward_prefetch(gsn_context* context, const Camera& cam, const int64_t* tile_offsets) {
    const int count = context->projected_gaussians;
    BackwardBuffers buffers;
    GSN_TRY(backward_buffers(context, cam, count, &buffers));  // allocated by the fork: no growth here
    GSN_HIP(hipStreamWaitEvent(context->aux_stream, context->fork_event, 0));
    GSN_TRY(launch_backward_preparation(context->aux_stream, cam, count, tile_offsets, context->projected_means,
        context->projected_conics, context->projected_colors, context->projected_rects,
        context->projected_record_offsets, buffers));
    GSN_HIP(hipEventRecord(context->join_event, context->aux_stream));
    context->prefetch_pending = true;
    BackwardPrefetch& queued = context->prefetch;
    queued.means = context->projected_means;
    queued.conics = context->projected_conics;
    queued.colors = context->projected_colors;
    queued.rects = context->projected_rects;
    queued.offsets = context->projected_record_offsets;
    queued.tile_offsets = tile_offsets;
    queued.count = count;
    queued.tiles_x = cam.tile_count_x;
    queued.tiles_y = cam.tile_count_y;
    queued.generation = context->projection_generation;
    return GSN_OK;
}

int launch_splat_backward(gsn_context* context, const Camera& cam, int count, const int* ids,
    const int64_t* tile_offsets, const float* means, const float* conics, const float* colors, const int* rects,
    const int* record_offsets, const float* image, const float* image_gradients, float* partials, int record_count,
    const uint8_t** flags, uint8_t* epoch, const int** gaussian_order, const StoredState* stored) {
    if (cam.tile_count_x > 0xffff || cam.tile_count_y > 0x7fff) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "raster backward: tile grid exceeds 65535 x 32767");
    }
    int64_t tiles = static_cast<int64_t>(cam.tile_count_x) * cam.tile_count_y;
    // Work queued by an armed gsn_render_tiles is joined first, and reused when it was computed
    // from these buffers of the current projection.
    const BackwardPrefetch& queued = context->prefetch;
    const bool prefetched = context->prefetch_pending && queued.generation == context->projection_generation &&
        queued.means == means && queued.conics == conics && queued.colors == colors && queued.rects == rects &&
        queued.offsets == record_offsets && queued.tile_offsets == tile_offsets && queued.count == count &&
        queued.tiles_x == cam.tile_count_x && queued.tiles_y == cam.tile_count_y;
    GSN_TRY(join_backward_prefetch(context));
    context->prefetch = BackwardPrefetch{};
    BackwardBuffers buffers;
    GSN_TRY(backward_buffers(context, cam, count, &buffers));
    PackedSplat* splats = buffers.splats;
    int* order = buffers.order;
    int* groups_order = buffers.groups_order;
    // One flag byte per record, read 16 at a time by the per-Gaussian pass (scratch slots are
    // allocated in multiples of 256 bytes, so an aligned 16-byte read below the record count
    // stays inside the allocation).
    uint8_t* record_flags = static_cast<uint8_t*>(scratch(context, kSlotRecordFlags, static_cast<size_t>(record_count)));
    if (!splats || !order || !groups_order || !record_flags) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    // The epoch runs 1..255. After a wrap, or in a newly allocated buffer, every byte is cleared
    // first, so that a stale byte never equals the current epoch.
    uint8_t next = static_cast<uint8_t>(context->record_epoch + 1);
    const size_t flag_bytes = context->slot_bytes[kSlotRecordFlags];
    if (next == 0 || context->record_flags_cleared != record_flags ||
        context->record_flags_cleared_bytes != flag_bytes) {
        int64_t words = static_cast<int64_t>(flag_bytes / 4);
        clear_flags_kernel<<<grid_for(words, 256), 256, 0, context->stream>>>(reinterpret_cast<uint32_t*>(record_flags),
            words);
        GSN_TRY(check_launch("clear_flags_kernel"));
        context->record_flags_cleared = record_flags;
        context->record_flags_cleared_bytes = flag_bytes;
        next = 1;
    }
    context->record_epoch = next;
    *flags = record_flags;
    *epoch = next;
    StageScope scope(context, kStageRasterBackward);
    if (prefetched) {
        ++context->prefetch_reused;
    }
    else {
        GSN_TRY(launch_backward_preparation(context->stream, cam, count, tile_offsets, means, conics, colors, rects,
            record_offsets, buffers));
    }
    *gaussian_order = groups_order;
    if (stored) {
        raster_backward_stored_kernel<<<static_cast<unsigned>(tiles), 64, 0, context->stream>>>(ids, tile_offsets,
            splats, order, cam, stored->final_transmittance, stored->last_contributor, image_gradients,
            stored->background[0], stored->background[1], stored->background[2], partials, record_flags, next,
            record_count);
        return check_launch("raster_backward_stored_kernel");
    }
    raster_backward_kernel<<<static_cast<unsigned>(tiles), 64, 0, context->stream>>>(ids, tile_offsets, splats, order,
        cam, image, image_gradients, partials, record_flags, next, record_count);
    return check_launch("raster_backward_kernel");
}

} // namespace gsn

namespace gsn {
namespace {

__global__ __launch_bounds__(256) void band_mask_selftest_kernel(const float* splats, const int* origins, int count,
    int* masks) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    const float* s = splats + static_cast<int64_t>(i) * 6;
    masks[i] = band_mask(s[0], s[1], make_float4(s[2], s[3], s[4], s[5]), origins[2 * i], origins[2 * i + 1]);
}

} // namespace
} // namespace gsn

// Diagnostic: the rasterisers' band mask on `count` (splat, tile) cases.
extern "C" int gsn_selftest_band_mask(gsn_context* context, const float* splats, const int* origins, int count,
    int* masks) {
    using namespace gsn;
    if (!context || count < 0 || (count > 0 && (!splats || !origins || !masks))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_selftest_band_mask: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    band_mask_selftest_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(splats, origins, count, masks);
    return check_launch("band_mask_selftest_kernel");
}

// Diagnostic: runs the wave reduction on one wave. input float[64][12] -> output float[64][3].
extern "C" int gsn_selftest_wave_sum12(gsn_context* context, const float* input, float* output) {
    using namespace gsn;
    if (!context || !input || !output) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_selftest_wave_sum12: null argument");
    }
    wave_sum12_selftest_kernel<<<1, 64, 0, context->stream>>>(input, output);
    return check_launch("wave_sum12_selftest_kernel");
}
