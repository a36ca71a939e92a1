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

// One 4096-key radix tile per workgroup (512 threads, eight Gaussians each, k * 512 apart so each
// access stays coalesced; loads issued before the stores). The workgroup also counts its keys'
// low eight bits, the first depth-sort pass's histogram column, so that pass needs no histogram
// kernel.
constexpr int kDepthKeyThreads = 512;
constexpr int kDepthKeyItems = kRadixTile / kDepthKeyThreads;

__global__ __launch_bounds__(kDepthKeyThreads) void depth_keys_kernel(const float* depths, const int* record_offsets,
    uint32_t* keys, uint32_t* ids, int count, int tiles, int* histogram) {
    __shared__ int s_digits[256];
    const int t = threadIdx.x;
    if (t < 256) {
        s_digits[t] = 0;
    }
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kRadixTile + t;
    float depth[kDepthKeyItems];
    int begin[kDepthKeyItems], end[kDepthKeyItems];
#pragma unroll
    for (int k = 0; k < kDepthKeyItems; ++k) {
        const int64_t i = first + k * kDepthKeyThreads;
        if (i < count) {
            depth[k] = depths[i];
            begin[k] = record_offsets[i];
            end[k] = record_offsets[i + 1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kDepthKeyItems; ++k) {
        const int64_t i = first + k * kDepthKeyThreads;
        if (i < count) {
            const uint32_t key = end[k] > begin[k] ? float_to_bits(depth[k]) : 0xffffffffu;
            keys[i] = key;
            ids[i] = static_cast<uint32_t>(i);
            atomicAdd(&s_digits[key & 0xffu], 1);
        }
    }
    __syncthreads();
    if (t < 256) {
        histogram[static_cast<int64_t>(t) * tiles + blockIdx.x] = s_digits[t];
    }
}

// Record counts and tile rectangles in depth order: the one gather by Gaussian id of the tile
// lists. A visible Gaussian's record count is its rectangle's area; invisible Gaussians (sorted
// last, with the key 0xffffffff that no finite depth has) get none, and their rectangles, which
// the projection leaves unwritten, are not read.
// One kScanBlock-entry block of [0, count] per workgroup (512 threads, eight entries each, k * 512
// apart): the workgroup also sums its counts into block_totals[block], the first half of the
// counts' exclusive scan.
constexpr int kOrderThreads = 512;
constexpr int kOrderItems = kScanBlock / kOrderThreads;

__global__ __launch_bounds__(kOrderThreads) void depth_order_counts_kernel(const uint32_t* sorted_ids,
    const uint32_t* sorted_keys, const int4* tile_rects, int* counts, int4* ordered_rects, int count,
    int* block_totals) {
    __shared__ int s_wave_sums[kOrderThreads / 64];
    const int t = threadIdx.x;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kScanBlock + t;
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kOrderItems; ++j) {
        const int64_t k = first + j * kOrderThreads;
        if (k < count) {
            int4 rect = make_int4(0, 0, 0, 0);
            if (sorted_keys[k] != 0xffffffffu) {
                rect = tile_rects[min(sorted_ids[k], static_cast<uint32_t>(count - 1))];
            }
            ordered_rects[k] = rect;
            const int records = max(rect.z - rect.x, 0) * max(rect.w - rect.y, 0);
            counts[k] = records;
            sum += records;
        }
        else if (k == count) {
            counts[k] = 0;
        }
    }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        sum += __shfl_xor(sum, offset, 64);
    }
    if ((t & 63) == 0) {
        s_wave_sums[t >> 6] = sum;
    }
    __syncthreads();
    if (t == 0) {
        int total = 0;
#pragma unroll
        for (int w = 0; w < kOrderThreads / 64; ++w) {
            total += s_wave_sums[w];
        }
        block_totals[blockIdx.x] = total;
    }
}

// Records are emitted in blocks of kRadixTile consecutive positions, one workgroup per block, so
// every workgroup writes the same amount whatever the footprints (a near-camera Gaussian can own
// thousands of records, a distant one a single record). A workgroup first needs the Gaussians
// (in depth order) that own its positions: emit_owners_kernel stores, for every block, the
// depth-order index of the Gaussian owning the block's first record, and in entry `blocks` the
// owner of the last record, so block b's owners are the contiguous range [owners[b], owners[b+1]].
constexpr int kEmitThreads = 512;
constexpr int kEmitItems = kRadixTile / kEmitThreads;
constexpr int kEmitOwners = kRadixTile + 1;

__global__ __launch_bounds__(256) void emit_owners_kernel(const int* depth_offsets, int count, int record_count,
    int blocks, int* owners) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) {
        return;
    }
    int64_t begin = max(depth_offsets[k], 0);
    int64_t end = min(depth_offsets[k + 1], record_count);
    if (end <= begin) {
        return;
    }
    for (int64_t b = (begin + kRadixTile - 1) / kRadixTile; b < blocks && b * kRadixTile < end; ++b) {
        owners[b] = k;
    }
    if (begin <= record_count - 1 && record_count - 1 < end) {
        owners[blocks] = k;
    }
}

// One workgroup per block of kRadixTile records. Owner data (record start, id, rectangle) is
// staged in LDS, every position is mapped to its owner by a max-scan over the positions where
// owners start, and consecutive lanes then write consecutive records: (row-major tile inside the
// owner's rectangle, Gaussian id). The workgroup also counts the low `digit_bits` bits of its
// tile ids, which is the first pass's histogram column of the radix sort that follows (the
// block is exactly one radix tile).
__global__ __launch_bounds__(kEmitThreads) void emit_records_kernel(const uint32_t* __restrict__ sorted_ids,
    const int* __restrict__ depth_offsets, const int* __restrict__ owners, const int4* __restrict__ ordered_rects,
    uint32_t* __restrict__ record_tiles, uint32_t* __restrict__ record_ids, int count, int tile_count_x,
    int record_count, int blocks, int digit_bits, int* __restrict__ histogram) {
    __shared__ int s_owner[kRadixTile];
    __shared__ int s_start[kEmitOwners];
    __shared__ uint32_t s_id[kEmitOwners];
    __shared__ uint32_t s_corner[kEmitOwners];  // x0 | y0 << 16
    __shared__ uint16_t s_width[kEmitOwners];
    __shared__ int s_digits[256];
    __shared__ int s_wave_max[kEmitThreads / 64];
    const int t = threadIdx.x;
    const int block = blockIdx.x;
    const int64_t base = static_cast<int64_t>(block) * kRadixTile;
    const int valid = static_cast<int>(min(static_cast<int64_t>(kRadixTile), record_count - base));
    const uint32_t digit_mask = (1u << digit_bits) - 1u;
    if (t < 256) {
        s_digits[t] = 0;
    }
    const int first = min(max(owners[block], 0), count - 1);
    const int last = min(max(owners[block + 1], first), count - 1);
    const int owner_count = last - first + 1;
    if (owner_count <= kEmitOwners) {
        for (int i = t; i < kRadixTile; i += kEmitThreads) {
            s_owner[i] = 0;
        }
        __syncthreads();
        for (int o = t; o < owner_count; o += kEmitThreads) {
            int k = first + o;
            int start = depth_offsets[k];
            uint32_t id = min(sorted_ids[k], static_cast<uint32_t>(count - 1));
            int4 rect = ordered_rects[k];
            s_start[o] = start;
            s_id[o] = id;
            s_corner[o] = static_cast<uint32_t>(rect.x & 0xffff) | (static_cast<uint32_t>(rect.y) << 16);
            s_width[o] = static_cast<uint16_t>(max(rect.z - rect.x, 1));
            int64_t position = static_cast<int64_t>(start) - base;
            if (position < kRadixTile) {
                // The block's first owner starts at or before the block; later owners start inside
                // it, and of several owners starting at one position (owners without records) the
                // last is the one whose records follow.
                atomicMax(&s_owner[max(position, int64_t{0})], o);
            }
        }
        __syncthreads();
        // Inclusive max-scan over the positions: thread t owns positions [kEmitItems t, kEmitItems (t + 1)).
        static_assert(kEmitItems % 4 == 0, "positions are read as int4");
        int v[kEmitItems];
#pragma unroll
        for (int q = 0; q < kEmitItems / 4; ++q) {
            const int4 quad = *reinterpret_cast<const int4*>(&s_owner[kEmitItems * t + 4 * q]);
            v[4 * q] = quad.x;
            v[4 * q + 1] = quad.y;
            v[4 * q + 2] = quad.z;
            v[4 * q + 3] = quad.w;
        }
#pragma unroll
        for (int i = 1; i < kEmitItems; ++i) {
            v[i] = max(v[i], v[i - 1]);
        }
        const int lane = t & 63;
        const int wave = t >> 6;
        int scan = v[kEmitItems - 1];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int other = __shfl_up(scan, d, 64);
            if (lane >= d) {
                scan = max(scan, other);
            }
        }
        if (lane == 63) {
            s_wave_max[wave] = scan;
        }
        int carry = __shfl_up(scan, 1, 64);
        if (lane == 0) {
            carry = 0;
        }
        __syncthreads();
        for (int w = 0; w < wave; ++w) {
            carry = max(carry, s_wave_max[w]);
        }
#pragma unroll
        for (int q = 0; q < kEmitItems / 4; ++q) {
            *reinterpret_cast<int4*>(&s_owner[kEmitItems * t + 4 * q]) = make_int4(max(v[4 * q], carry),
                max(v[4 * q + 1], carry), max(v[4 * q + 2], carry), max(v[4 * q + 3], carry));
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kEmitItems; ++j) {
            int p = j * kEmitThreads + t;
            if (p < valid) {
                int o = s_owner[p];
                int local = static_cast<int>(base + p - s_start[o]);
                int width = s_width[o];
                int row = local / width;
                int column = local - row * width;
                uint32_t corner = s_corner[o];
                uint32_t tile = static_cast<uint32_t>(((corner >> 16) + row) * tile_count_x + (corner & 0xffffu) + column);
                // Written once and read once by the next sort pass: non-temporal stores keep the
                // 240 MB of records (30M at 1080p) from displacing cached data (emit + tile sort
                // 0.506 -> 0.473 ms).
                __builtin_nontemporal_store(tile, record_tiles + base + p);
                __builtin_nontemporal_store(s_id[o], record_ids + base + p);
                atomicAdd(&s_digits[tile & digit_mask], 1);
            }
        }
    }
    else {
        // Only reachable when Gaussians without records sit between ones with records (the
        // projection never produces that): find each position's owner by binary search.
        __syncthreads();
        for (int j = 0; j < kEmitItems; ++j) {
            int p = j * kEmitThreads + t;
            if (p >= valid) {
                continue;
            }
            int r = static_cast<int>(base + p);
            int low = 0, high = count;
            while (high - low > 1) {
                int middle = (low + high) >> 1;
                if (depth_offsets[middle] <= r) {
                    low = middle;
                }
                else {
                    high = middle;
                }
            }
            uint32_t id = min(sorted_ids[low], static_cast<uint32_t>(count - 1));
            int4 rect = ordered_rects[low];
            int width = max(rect.z - rect.x, 1);
            int local = r - depth_offsets[low];
            int row = local / width;
            int column = local - row * width;
            uint32_t tile = static_cast<uint32_t>((rect.y + row) * tile_count_x + rect.x + column);
            record_tiles[r] = tile;
            record_ids[r] = id;
            atomicAdd(&s_digits[tile & digit_mask], 1);
        }
    }
    __syncthreads();
    if (t < (1 << digit_bits)) {
        histogram[static_cast<int64_t>(t) * blocks + block] = s_digits[t];
    }
}

// Four records per thread (one 16-byte load when the run is whole): a tile's range starts at
// the first record whose tile differs from its predecessor's.
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t* sorted_tiles, int64_t* offsets, int records,
    int tile_count) {
    const int64_t first = 4 * (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x);
    if (first >= records) {
        return;
    }
    int tiles[4];
    if (first + 4 <= records) {
        const uint4 quad = *reinterpret_cast<const uint4*>(sorted_tiles + first);
        tiles[0] = static_cast<int>(quad.x);
        tiles[1] = static_cast<int>(quad.y);
        tiles[2] = static_cast<int>(quad.z);
        tiles[3] = static_cast<int>(quad.w);
    }
    else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            tiles[i] = first + i < records ? static_cast<int>(sorted_tiles[first + i]) : tile_count;
        }
    }
    // Tile ids come from rectangles clamped to the grid; the clamp keeps writes inside
    // offsets[0..T] even for rectangles produced with a different camera.
    int previous = first == 0 ? -1 : min(max(static_cast<int>(sorted_tiles[first - 1]), 0), tile_count);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t r = first + i;
        if (r >= records) {
            break;
        }
        const int tile = min(max(tiles[i], 0), tile_count);
        for (int t = previous + 1; t <= tile; ++t) {
            offsets[t] = r;
        }
        previous = tile;
        if (r == records - 1) {
            for (int t = tile + 1; t <= tile_count; ++t) {
                offsets[t] = records;
            }
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

// Stable order of the N Gaussians by depth bits (visible ones first) into the depth-order slots:
// sorted ids and sorted keys.
int enqueue_depth_order(gsn_context* context, const float* depths, const int* record_offsets, int count) {
    size_t n = static_cast<size_t>(count);
    uint32_t* depth_keys = static_cast<uint32_t*>(scratch(context, kSlotDepthKeys, n * 4));
    uint32_t* depth_ids = static_cast<uint32_t*>(scratch(context, kSlotDepthIds, n * 4));
    uint32_t* sorted_ids = static_cast<uint32_t*>(scratch(context, kSlotDepthSortedIds, n * 4));
    uint32_t* sorted_keys = static_cast<uint32_t*>(scratch(context, kSlotDepthSortedKeys, n * 4));
    if (!depth_keys || !depth_ids || !sorted_ids || !sorted_keys) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    int* histogram = radix_histogram_buffer(context, count);
    if (!histogram) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    static_assert(kDepthKeyThreads * kDepthKeyItems == kRadixTile, "one radix tile per workgroup");
    const int tiles = static_cast<int>((static_cast<int64_t>(count) + kRadixTile - 1) / kRadixTile);
    StageScope scope(context, kStageDepthSort);
    // 32 key bits: the first pass takes radix_first_digit_bits(32) = 8 bits, the low byte counted here.
    depth_keys_kernel<<<tiles, kDepthKeyThreads, 0, context->stream>>>(depths, record_offsets, depth_keys, depth_ids,
        count, tiles, histogram);
    GSN_TRY(check_launch("depth_keys_kernel"));
    return radix_sort_pairs(context, depth_keys, depth_ids, count, 32, sorted_keys, sorted_ids, true);
}

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
    GSN_HIP(hipEventRecord(context->readback_event, context->stream));
    // The depth order of the last projection does not need R: queue it behind the copy, so the
    // GPU sorts while the caller reads R and queues the kernels that do.
    const bool prepare = count > 0 && context->projected_depths && context->projected_offsets == record_offsets &&
                         context->projected_count == count;
    const float* depths = context->projected_depths;
    context->projected_depths = nullptr;
    context->projected_offsets = nullptr;
    context->projected_count = -1;
    if (prepare) {
        GSN_TRY(enqueue_depth_order(context, depths, record_offsets, count));
        context->ordered_depths = depths;
        context->ordered_offsets = record_offsets;
        context->ordered_count = count;
    }
    GSN_HIP(hipEventSynchronize(context->readback_event));
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
    if (camera->tile_count_x > 0xffff || camera->tile_count_y > 0xffff) {
        return fail(GSN_ERROR_CAPACITY, "gsn_build_tile_lists: more than 65535 tiles along an axis");
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
    const bool ordered = context->ordered_depths == depths && context->ordered_offsets == record_offsets &&
                         context->ordered_count == count;
    context->ordered_depths = nullptr;
    context->ordered_offsets = nullptr;
    context->ordered_count = -1;
    if (!ordered) {
        GSN_TRY(enqueue_depth_order(context, depths, record_offsets, count));
    }
    uint32_t* sorted_ids = static_cast<uint32_t*>(scratch(context, kSlotDepthSortedIds, n * 4));
    uint32_t* depth_sorted_keys = static_cast<uint32_t*>(scratch(context, kSlotDepthSortedKeys, n * 4));
    int4* ordered_rects = static_cast<int4*>(scratch(context, kSlotDepthRects, n * sizeof(int4)));
    int* depth_counts = static_cast<int*>(scratch(context, kSlotDepthCounts, (n + 1) * 4));
    uint32_t* record_tiles = static_cast<uint32_t*>(scratch(context, kSlotRecordTiles, r * 4));
    uint32_t* record_ids = static_cast<uint32_t*>(scratch(context, kSlotRecordIds, r * 4));
    if (!sorted_ids || !depth_sorted_keys || !ordered_rects || !depth_counts || !record_tiles || !record_ids) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    int tile_bits = 0;
    for (unsigned t = static_cast<unsigned>(tile_count - 1); t; t >>= 1) {
        ++tile_bits;
    }
    const int blocks = static_cast<int>((static_cast<int64_t>(record_count) + kRadixTile - 1) / kRadixTile);
    uint32_t* sorted_tiles = static_cast<uint32_t*>(scratch(context, kSlotSortKeysB, r * 4));
    int* block_owners = static_cast<int*>(scratch(context, kSlotEmitOwners, (static_cast<size_t>(blocks) + 1) * 4));
    int* histogram = radix_histogram_buffer(context, record_count);
    const int count_blocks = static_cast<int>((static_cast<int64_t>(count) + kScanBlock) / kScanBlock);  // count + 1 entries
    int* count_totals = static_cast<int*>(scratch(context, kSlotDepthCountTotals, static_cast<size_t>(count_blocks) * 4));
    if (!sorted_tiles || !block_owners || !histogram || !count_totals) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    {
        StageScope scope(context, kStageEmit);
        depth_order_counts_kernel<<<count_blocks, kOrderThreads, 0, stream>>>(sorted_ids, depth_sorted_keys,
            reinterpret_cast<const int4*>(tile_rects), depth_counts, ordered_rects, count, count_totals);
        GSN_TRY(check_launch("depth_order_counts_kernel"));
        GSN_TRY(exclusive_scan_int_from_totals(context, depth_counts, depth_counts, static_cast<int64_t>(count) + 1,
            count_totals));
        emit_owners_kernel<<<grid_for(count, 256), 256, 0, stream>>>(depth_counts, count, record_count, blocks,
            block_owners);
        GSN_TRY(check_launch("emit_owners_kernel"));
        // With a single tile (tile_bits == 0) the histogram is not used; one digit bit keeps the
        // kernel's arguments in range.
        emit_records_kernel<<<blocks, kEmitThreads, 0, stream>>>(sorted_ids, depth_counts, block_owners,
            ordered_rects, record_tiles, record_ids, count, camera->tile_count_x,
            record_count, blocks, radix_first_digit_bits(tile_bits), histogram);
        GSN_TRY(check_launch("emit_records_kernel"));
    }
    {
        StageScope scope(context, kStageTileSort);
        if (tile_bits == 0) {
            GSN_HIP(hipMemcpyAsync(tile_gaussian_ids, record_ids, r * 4, hipMemcpyDeviceToDevice, stream));
            GSN_HIP(hipMemcpyAsync(sorted_tiles, record_tiles, r * 4, hipMemcpyDeviceToDevice, stream));
        }
        else {
            GSN_TRY(radix_sort_pairs(context, record_tiles, record_ids, record_count, tile_bits, sorted_tiles,
                reinterpret_cast<uint32_t*>(tile_gaussian_ids), true));
        }
    }
    // The ranges kernel writes every entry of tile_offsets[0..T] when R > 0.
    StageScope scope(context, kStageRanges);
    tile_ranges_kernel<<<grid_for((static_cast<int64_t>(record_count) + 3) / 4, 256), 256, 0, stream>>>(sorted_tiles,
        tile_offsets, record_count, tile_count);
    GSN_TRY(check_launch("tile_ranges_kernel"));
    if (sorted_keys) {
        sorted_keys_kernel<<<grid_for(record_count, 256), 256, 0, stream>>>(sorted_tiles, tile_gaussian_ids, depths,
            sorted_keys, record_count);
        GSN_TRY(check_launch("sorted_keys_kernel"));
    }
    return GSN_OK;
}
