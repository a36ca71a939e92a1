// Internal declarations shared by the HIP translation units of libgsn_hip.so.
#pragma once

#include "gsn_common.h"
#include "../../include/gsn_hip.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>

// Scratch slots owned by a context. Each slot grows on demand and is reused across calls.
enum ScratchSlot {
    kSlotSortKeysA = 0,
    kSlotSortKeysB,
    kSlotSortValuesA,
    kSlotSortValuesB,
    kSlotSortHistogram,
    kSlotScanBlocks,
    kSlotDepthKeys,
    kSlotDepthIds,
    kSlotDepthOffsets,
    kSlotDepthSortedIds,
    kSlotDepthCounts,
    kSlotRecordTiles,
    kSlotRecordIds,
    kSlotPartials,
    kSlotReduce,
    kSlotKnnKeys,
    kSlotKnnIds,
    kSlotKnnNodes,
    kSlotKnnPoints,
    kSlotMisc,
    kSlotSplats,
    kSlotRecordFlags,
    kSlotTileOrder,
    kSlotEmitOwners,
    kSlotDepthSortedKeys,
    kSlotDepthRects,
    kSlotGaussianOrder,
    kSlotQuantizedError,
    kSlotKnnOrder,
    kSlotDepthCountTotals,
    kSlotCount
};

// Stages timed with HIP events when a context's stage timing is enabled (order = gsn_stage).
enum Stage {
    kStageProject = 0,
    kStageScan,
    kStageDepthSort,
    kStageEmit,
    kStageTileSort,
    kStageRanges,
    kStageRasterForward,
    kStageImageGradients,
    kStageRasterBackward,
    kStageGaussianGradients,
    kStageAdam,
    kStageDensity,
    kStageKnn,
    kStageAccumulate,
    kStageCount
};

// The backward's splat copy and schedules queued on the auxiliary stream by an armed
// gsn_render_tiles, and the buffers they were computed from.
struct BackwardPrefetch {
    const float* means = nullptr;
    const float* conics = nullptr;
    const float* colors = nullptr;
    const int* rects = nullptr;
    const int* offsets = nullptr;
    const int64_t* tile_offsets = nullptr;
    int count = -1;
    int tiles_x = 0;
    int tiles_y = 0;
    unsigned generation = 0;
};

struct gsn_context {
    hipStream_t stream = nullptr;
    int device = 0;
    int cu_count = 0;
    void* slots[kSlotCount] = {};
    size_t slot_bytes[kSlotCount] = {};
    void* pinned = nullptr;  // small pinned host buffer for scalar readbacks
    // Epoch stamp of the backward's record flags (1..255) and the flag buffer (address and size:
    // a regrown buffer may reuse the address) it was last cleared in: a byte equals the current
    // epoch only if this step's backward reached that record.
    uint8_t record_epoch = 0;
    void* record_flags_cleared = nullptr;
    size_t record_flags_cleared_bytes = 0;
    // Depth order queued ahead of gsn_build_tile_lists: the buffers of the last projection, and
    // those whose order gsn_scan_record_offsets has queued (consumed by the next tile-list
    // build). readback_event marks the record-count copy.
    const float* projected_depths = nullptr;
    const int* projected_offsets = nullptr;
    int projected_count = -1;
    // The last projection's other outputs, and a count of projections.
    const float* projected_means = nullptr;
    const float* projected_conics = nullptr;
    const float* projected_colors = nullptr;
    const int* projected_rects = nullptr;
    const int* projected_record_offsets = nullptr;  // kept when the depth-order fields are consumed
    int projected_gaussians = -1;
    unsigned projection_generation = 0;
    // Backward prefetch. gsn_context_prefetch_backward arms the next gsn_render_tiles of the last
    // projection: it queues the backward's splat packing and schedules on aux_stream behind
    // fork_event (the main stream's work so far) and records join_event. The gradient entry
    // points, the next projection and the next prefetch make the main stream wait for join_event
    // first; the gradient entry points reuse the queued work when their buffers are the prefetch's.
    bool prefetch_armed = false;
    bool prefetch_pending = false;
    unsigned long long prefetch_reused = 0;  // gradient calls that used queued work (diagnostic)
    BackwardPrefetch prefetch;
    hipStream_t aux_stream = nullptr;
    hipEvent_t fork_event = nullptr;
    hipEvent_t join_event = nullptr;
    const float* ordered_depths = nullptr;
    const int* ordered_offsets = nullptr;
    int ordered_count = -1;
    hipEvent_t readback_event = nullptr;
    // Stage timing: event pairs recorded around the launches of the stages in timing_mask.
    unsigned timing_mask = 0;
    hipEvent_t* events = nullptr;
    int* event_stage = nullptr;
    int event_capacity = 0;  // number of pairs
    int event_used = 0;
    // Raster pair counting (diagnostic): while enabled, the raster forward adds the number of
    // (pixel, splat) pairs it evaluates to *raster_pairs (device memory owned by the context).
    bool count_raster_pairs = false;
    unsigned long long* raster_pairs = nullptr;
    // Device address of the allocation-free metrics' file-scope array on this context's device
    // (resolved by gsn_context_create).
    void* static_scratch = nullptr;
};

namespace gsn {

// Records an event pair around the launches in its scope when timing is enabled. A scope made
// with kernel_events wraps exactly one kernel and records nothing itself: the kernel is launched
// with start_event() and stop_event() through hipExtLaunchKernel, whose dispatch records them
// (no marker packets of their own between the stage's kernel and its neighbours). Both are null
// when the stage is not timed.
class StageScope {
public:
    StageScope(gsn_context* context, Stage stage, bool kernel_events = false);
    ~StageScope();
    hipEvent_t start_event() const;
    hipEvent_t stop_event() const;

private:
    gsn_context* context_;
    int pair_;
    bool kernel_events_;
};

} // namespace gsn

namespace gsn {

// Records a failure message for gsn_last_error() and returns the status.
int fail(int status, const char* format, ...);
int check_hip(hipError_t error, const char* what);
int check_launch(const char* kernel);
// Returns a device buffer of at least `bytes` for `slot`, growing it if needed (no copy).
void* scratch(gsn_context* context, ScratchSlot slot, size_t bytes);

// Image metrics (ssim.hip): partial sums per 32x16 block and channel into block_sums
// (image_metric_blocks(width, height) x 3 doubles), then a fixed-order finish into out[4] (device).
int64_t image_metric_blocks(int width, int height);
int launch_image_metrics(gsn_context* context, int width, int height, const float* image, const uint8_t* target,
    double* block_sums, double* out);
// Device address of the allocation-free metrics' array on `device` (static_metrics.hip); the
// device must be current.
int resolve_static_scratch(int device, void** address);

// Device-wide exclusive prefix scan of int / int4-like arrays (scan.hip).
struct Int4 {
    int x, y, z, w;
};
int exclusive_scan_int(gsn_context* context, const int* input, int* output, int64_t n);
// The same scan when the producer of `input` has written the sum of each kScanBlock-element block
// (entries [b * kScanBlock, (b + 1) * kScanBlock) of [0, n)) to block_totals[b]: only the
// downsweep runs (beyond 2048 blocks the totals are ignored and the full scan runs).
constexpr int kScanBlock = 4096;
int exclusive_scan_int_from_totals(gsn_context* context, const int* input, int* output, int64_t n,
    const int* block_totals);
int exclusive_scan_int4(gsn_context* context, const Int4* input, Int4* output, int64_t n);

// Stable LSD radix sort of (uint32 key, uint32 value) pairs over key bits [0, bits)
// (radix_sort.hip). Input buffers are consumed. The result is written to keys_out/values_out
// (either may be the same buffer as an input only if it is not referenced by a later pass;
// callers pass distinct buffers). keys_out may be null when only values are needed.
// With first_histogram_ready the caller has already written the first pass's digit counts into
// radix_histogram_buffer(context, n): entry [bin * tiles + tile] counts the elements of
// [tile * kRadixTile, tile * kRadixTile + kRadixTile) whose low radix_first_digit_bits(bits) key
// bits equal bin, for every bin below 1 << radix_first_digit_bits(bits).
constexpr int kRadixTile = 4096;
int* radix_histogram_buffer(gsn_context* context, int64_t n);
int radix_first_digit_bits(int bits);
int radix_sort_pairs(gsn_context* context, uint32_t* keys, uint32_t* values, int64_t n,
    int bits, uint32_t* keys_out, uint32_t* values_out, bool first_histogram_ready = false);

// Floats per (tile, Gaussian) backward partial: dmean2D(2), dSigma2D(3), dlogit, dcolor(3),
// |dmean2D|(2), pad.
constexpr int kPartialFloats = 12;

// Gaussians per workgroup of the per-Gaussian gradient pass.
constexpr int kGradientGroup = 256;

// Per-pixel forward state consumed by the stored-state backward.
struct StoredState {
    const float* final_transmittance;
    const int* last_contributor;
    float background[3];
};

// Tile rasteriser backward writing per-record partials (raster_backward.hip): the recompute form
// when stored is null, else the back-to-front stored-state form. Only the records a tile reaches
// before all its pixels are done get a partial; for each of them the backward also stores the
// step's epoch into record_flags[slot] (*flags, *epoch are returned), so the per-Gaussian pass
// reads exactly the partials written in this step.
// Makes these buffers the context's last projection (project.hip): the tile-list build, the depth
// order queued by the record scan and the backward prefetch refer to it.
int record_projection(gsn_context* context, int count, float* means, float* conic_opacities, float* colors,
    float* depths, int* tile_rects, int* record_offsets);

// The backward prefetch of an armed gsn_render_tiles (raster_backward.hip), in two parts so that
// the forward kernel is launched in between: fork_backward_prefetch joins earlier queued work,
// allocates the buffers and records the fork event on the main stream (*forked: whether it did);
// queue_backward_prefetch then queues the splat copy and schedules of the last projection on the
// auxiliary stream behind that event.
int fork_backward_prefetch(gsn_context* context, const Camera& cam, bool* forked);
int queue_backward_prefetch(gsn_context* context, const Camera& cam, const int64_t* tile_offsets);
// Makes the main stream wait for queued prefetch work (nothing if none is pending).
int join_backward_prefetch(gsn_context* context);

int launch_splat_backward(gsn_context* context, const Camera& cam, int count, const int* ids,
    const int64_t* tile_offsets, const float* means, const float* conics, const float* colors, const int* rects,
    const int* record_offsets, const float* image, const float* image_gradients, float* partials, int record_count,
    const uint8_t** flags, uint8_t* epoch, const int** gaussian_order,
    const StoredState* stored = nullptr);

// True when every row of a [rows][count] float array starts on a 128-byte line: a wave's
// 256-byte segment of a row then covers two whole lines (the condition under which the streaming
// kernels use non-temporal accesses).
inline bool rows_line_aligned(const void* base, int64_t count) {
    return count % 32 == 0 && reinterpret_cast<uintptr_t>(base) % 128 == 0;
}

inline unsigned grid_for(int64_t n, int block) {
    return static_cast<unsigned>((n + block - 1) / block);
}

} // namespace gsn

#define GSN_TRY(expr)                 \
    do {                              \
        int gsn_status__ = (expr);    \
        if (gsn_status__ != GSN_OK) { \
            return gsn_status__;      \
        }                             \
    } while (0)

#define GSN_HIP(expr) GSN_TRY(::gsn::check_hip((expr), #expr))
