// Tile keys, stable (tile, depth) ordering and per-tile ranges.
//
// The reference (renderer.cu:223-271, 534-569) writes 64-bit keys (tile << 32 | depth bits)
// for every record in Gaussian-id order and radix-sorts all 32 + ceil(log2 T) key bits over R
// records. This path reaches the same order with far fewer passes over R:
//   1. stable sort of the N visible Gaussians by depth bits (4 x 8-bit passes over N),
//   2. records emitted in that depth order (row-major tiles inside each rectangle),
//   3. stable sort of the records by tile id only (ceil(log2 T) bits, two passes for T <= 65536).
// Within a tile the result is ordered by (depth bits, Gaussian id), exactly the order of a
// stable sort of the reference keys over id-ordered records (each Gaussian has at most one
// record per tile, and equal depth bits keep id order in both).
#include "gsn_internal.h"

namespace gsn {
namespace {

__global__ __launch_bounds__(256) void depth_keys_kernel(const float* depths, const int* record_offsets, uint32_t* keys,
    uint32_t* ids, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) {
        return;
    }
    bool visible = record_offsets[i + 1] > record_offsets[i];
    keys[i] = visible ? float_to_bits(depths[i]) : 0xffffffffu;
    ids[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(256) void depth_order_counts_kernel(const uint32_t* sorted_ids, const int* record_offsets,
    int* counts, int count) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > count) {
        return;
    }
    if (k == count) {
        counts[k] = 0;
        return;
    }
    int id = static_cast<int>(sorted_ids[k]);
    counts[k] = record_offsets[id + 1] - record_offsets[id];
}

// Eight lanes per Gaussian enumerate its rectangle row-major.
constexpr int kEmitLanes = 8;

__global__ __launch_bounds__(256) void emit_records_kernel(const uint32_t* sorted_ids, const int* depth_offsets,
    const int4* tile_rects, uint32_t* record_tiles, uint32_t* record_ids, int count, int tile_count_x,
    int record_count) {
    int k = (blockIdx.x * blockDim.x + threadIdx.x) / kEmitLanes;
    int lane = threadIdx.x % kEmitLanes;
    if (k >= count) {
        return;
    }
    int begin = depth_offsets[k];
    int total = depth_offsets[k + 1] - begin;
    if (total == 0) {
        return;
    }
    uint32_t id = sorted_ids[k];
    int4 rect = tile_rects[id];
    int width = rect.z - rect.x;
    // record_count bounds the writes even if record_offsets disagree with it.
    for (int local = lane; local < total && begin + local < record_count; local += kEmitLanes) {
        int row = local / width;
        int column = local - row * width;
        record_tiles[begin + local] = static_cast<uint32_t>((rect.y + row) * tile_count_x + rect.x + column);
        record_ids[begin + local] = id;
    }
}

__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t* sorted_tiles, int64_t* offsets, int records,
    int tile_count) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= records) {
        return;
    }
    // Tile ids come from rectangles clamped to the grid; min() keeps writes inside offsets[0..T]
    // even for rectangles produced with a different camera.
    int tile = min(static_cast<int>(sorted_tiles[r]), tile_count);
    int previous = r == 0 ? -1 : min(static_cast<int>(sorted_tiles[r - 1]), tile_count);
    if (tile != previous) {
        for (int t = previous + 1; t <= tile; ++t) {
            offsets[t] = r;
        }
    }
    if (r == records - 1) {
        for (int t = tile + 1; t <= tile_count; ++t) {
            offsets[t] = records;
        }
    }
}

__global__ __launch_bounds__(256) void sorted_keys_kernel(const uint32_t* sorted_tiles, const int* ids, const float* depths,
    uint64_t* keys, int records) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= records) {
        return;
    }
    keys[r] = (static_cast<uint64_t>(sorted_tiles[r]) << 32) | float_to_bits(depths[ids[r]]);
}

__global__ void zero_offsets_kernel(int64_t* offsets, int entries) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < entries) {
        offsets[i] = 0;
    }
}

} // namespace
} // namespace gsn

extern "C" int gsn_scan_record_offsets(gsn_context* context, int* record_offsets, int count, int* record_count) {
    using namespace gsn;
    if (!context || !record_offsets || count < 0 || count == 0x7fffffff || !record_count) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_scan_record_offsets: invalid arguments");
    }
    {
        StageScope scope(context, kStageScan);
        GSN_TRY(exclusive_scan_int(context, record_offsets, record_offsets, static_cast<int64_t>(count) + 1));
    }
    int* pinned = static_cast<int*>(context->pinned);
    GSN_HIP(hipMemcpyAsync(pinned, record_offsets + count, sizeof(int), hipMemcpyDeviceToHost, context->stream));
    GSN_HIP(hipStreamSynchronize(context->stream));
    *record_count = pinned[0];
    if (*record_count < 0) {
        return fail(GSN_ERROR_CAPACITY, "gsn_scan_record_offsets: record count overflowed int32");
    }
    return GSN_OK;
}

extern "C" int gsn_build_tile_lists(gsn_context* context, const gsn_camera* camera, int count, const int* tile_rects,
    const float* depths, const int* record_offsets, int record_count, int* tile_gaussian_ids, int64_t* tile_offsets,
    uint64_t* sorted_keys) {
    using namespace gsn;
    if (!context || !camera || count < 0 || record_count < 0 || !tile_offsets || !record_offsets) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_build_tile_lists: invalid arguments");
    }
    int64_t tile_count_long = static_cast<int64_t>(camera->tile_count_x) * camera->tile_count_y;
    if (camera->tile_count_x <= 0 || camera->tile_count_y <= 0 || tile_count_long >= 0x7fffffff) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_build_tile_lists: invalid tile grid");
    }
    int tile_count = static_cast<int>(tile_count_long);
    hipStream_t stream = context->stream;
    if (record_count == 0 || count == 0) {
        zero_offsets_kernel<<<grid_for(tile_count + 1, 256), 256, 0, stream>>>(tile_offsets, tile_count + 1);
        return check_launch("zero_offsets_kernel");
    }
    if (!tile_rects || !depths || !tile_gaussian_ids) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_build_tile_lists: null buffer");
    }
    size_t n = static_cast<size_t>(count);
    size_t r = static_cast<size_t>(record_count);
    uint32_t* depth_keys = static_cast<uint32_t*>(scratch(context, kSlotDepthKeys, n * 4));
    uint32_t* depth_ids = static_cast<uint32_t*>(scratch(context, kSlotDepthIds, n * 4));
    uint32_t* sorted_ids = static_cast<uint32_t*>(scratch(context, kSlotDepthSortedIds, n * 4));
    int* depth_counts = static_cast<int*>(scratch(context, kSlotDepthCounts, (n + 1) * 4));
    uint32_t* record_tiles = static_cast<uint32_t*>(scratch(context, kSlotRecordTiles, r * 4));
    uint32_t* record_ids = static_cast<uint32_t*>(scratch(context, kSlotRecordIds, r * 4));
    if (!depth_keys || !depth_ids || !sorted_ids || !depth_counts || !record_tiles || !record_ids) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    {
        StageScope scope(context, kStageDepthSort);
        depth_keys_kernel<<<grid_for(count, 256), 256, 0, stream>>>(depths, record_offsets, depth_keys, depth_ids,
            count);
        GSN_TRY(check_launch("depth_keys_kernel"));
        GSN_TRY(radix_sort_pairs(context, depth_keys, depth_ids, count, 32, nullptr, sorted_ids));
    }
    {
        StageScope scope(context, kStageEmit);
        depth_order_counts_kernel<<<grid_for(count + 1, 256), 256, 0, stream>>>(sorted_ids, record_offsets,
            depth_counts, count);
        GSN_TRY(check_launch("depth_order_counts_kernel"));
        GSN_TRY(exclusive_scan_int(context, depth_counts, depth_counts, static_cast<int64_t>(count) + 1));
        emit_records_kernel<<<grid_for(static_cast<int64_t>(count) * kEmitLanes, 256), 256, 0, stream>>>(sorted_ids,
            depth_counts, reinterpret_cast<const int4*>(tile_rects), record_tiles, record_ids, count,
            camera->tile_count_x, record_count);
        GSN_TRY(check_launch("emit_records_kernel"));
    }
    int tile_bits = 0;
    for (unsigned t = static_cast<unsigned>(tile_count - 1); t; t >>= 1) {
        ++tile_bits;
    }
    uint32_t* sorted_tiles = static_cast<uint32_t*>(scratch(context, kSlotSortKeysB, r * 4));
    if (!sorted_tiles) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    {
        StageScope scope(context, kStageTileSort);
        if (tile_bits == 0) {
            GSN_HIP(hipMemcpyAsync(tile_gaussian_ids, record_ids, r * 4, hipMemcpyDeviceToDevice, stream));
            GSN_HIP(hipMemcpyAsync(sorted_tiles, record_tiles, r * 4, hipMemcpyDeviceToDevice, stream));
        }
        else {
            GSN_TRY(radix_sort_pairs(context, record_tiles, record_ids, record_count, tile_bits, sorted_tiles,
                reinterpret_cast<uint32_t*>(tile_gaussian_ids)));
        }
    }
    // The ranges kernel writes every entry of tile_offsets[0..T] when R > 0.
    StageScope scope(context, kStageRanges);
    tile_ranges_kernel<<<grid_for(record_count, 256), 256, 0, stream>>>(sorted_tiles, tile_offsets, record_count,
        tile_count);
    GSN_TRY(check_launch("tile_ranges_kernel"));
    if (sorted_keys) {
        sorted_keys_kernel<<<grid_for(record_count, 256), 256, 0, stream>>>(sorted_tiles, tile_gaussian_ids, depths,
            sorted_keys, record_count);
        GSN_TRY(check_launch("sorted_keys_kernel"));
    }
    return GSN_OK;
}
