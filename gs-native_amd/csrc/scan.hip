// Device-wide exclusive prefix scan: per-block totals, then a downsweep in which every block adds
// the totals of the blocks before it (summed directly up to 2048 blocks, else from a recursive
// scan of the totals).
// Used for record offsets (int[N+1]), depth-order record counts and density decisions (int4[N+1]).
#include "gsn_internal.h"

namespace gsn {
namespace {

// 512 threads x 8 items per 4096-element block (against 256 x 16: about 17 us less per training
// step over the scans of record offsets, depth-order counts and radix histograms).
constexpr int kScanThreads = 512;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
static_assert(kScanTile == kScanBlock, "block totals written by producers cover kScanBlock entries");

__device__ __forceinline__ int add(int a, int b) {
    return a + b;
}

__device__ __forceinline__ Int4 add(Int4 a, Int4 b) {
    return Int4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
}

template<typename T>
__device__ __forceinline__ T zero_value();

template<>
__device__ __forceinline__ int zero_value<int>() {
    return 0;
}

template<>
__device__ __forceinline__ Int4 zero_value<Int4>() {
    return Int4{0, 0, 0, 0};
}

__device__ __forceinline__ int shuffle_up(int v, int d) {
    return __shfl_up(v, d, 64);
}

__device__ __forceinline__ Int4 shuffle_up(Int4 v, int d) {
    return Int4{__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64), __shfl_up(v.z, d, 64), __shfl_up(v.w, d, 64)};
}

// Inclusive scan across the 64 lanes of a wave.
template<typename T>
__device__ __forceinline__ T wave_inclusive_scan(T value) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T other = shuffle_up(value, d);
        if (lane >= d) {
            value = add(value, other);
        }
    }
    return value;
}

// Exclusive scan of one value per thread across the block (wave scans, then the four wave
// totals through LDS); the block total is returned via *total. Called once per kernel.
template<typename T>
__device__ __forceinline__ T block_exclusive_scan(T value, T* total) {
    __shared__ T wave_totals[kScanThreads / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    T inclusive = wave_inclusive_scan(value);
    if (lane == 63) {
        wave_totals[wave] = inclusive;
    }
    T exclusive = shuffle_up(inclusive, 1);
    if (lane == 0) {
        exclusive = zero_value<T>();
    }
    __syncthreads();
    T before = zero_value<T>();
    T all = zero_value<T>();
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        T x = wave_totals[w];
        if (w < wave) {
            before = add(before, x);
        }
        all = add(all, x);
    }
    *total = all;
    return add(before, exclusive);
}

// Sum across the block, returned to every thread. Called once per kernel.
template<typename T>
__device__ __forceinline__ T block_sum(T value) {
    T total;
    block_exclusive_scan(value, &total);
    return total;
}

// Per-block totals of kScanTile consecutive inputs (coalesced loads: order does not matter for
// the sum).
template<typename T>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const T* input, T* block_totals, int64_t n) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
    T sum = zero_value<T>();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t index = base + i * kScanThreads + threadIdx.x;
        if (index < n) {
            sum = add(sum, input[index]);
        }
    }
    T total = block_sum(sum);
    if (threadIdx.x == 0) {
        block_totals[blockIdx.x] = total;
    }
}

// Loads / stores a thread's kScanItems consecutive values; int uses 16-byte accesses when the
// pointer allows.
__device__ __forceinline__ void load_items(const int* input, int64_t base, int64_t n, bool vector, int (&items)[kScanItems]) {
    if (vector && base + kScanItems <= n) {
        const int4* source = reinterpret_cast<const int4*>(input + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            int4 v = source[q];
            items[4 * q] = v.x;
            items[4 * q + 1] = v.y;
            items[4 * q + 2] = v.z;
            items[4 * q + 3] = v.w;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        items[i] = base + i < n ? input[base + i] : 0;
    }
}

__device__ __forceinline__ void store_items(int* output, int64_t base, int64_t n, bool vector, const int (&items)[kScanItems]) {
    if (vector && base + kScanItems <= n) {
        int4* destination = reinterpret_cast<int4*>(output + base);
#pragma unroll
        for (int q = 0; q < kScanItems / 4; ++q) {
            destination[q] = make_int4(items[4 * q], items[4 * q + 1], items[4 * q + 2], items[4 * q + 3]);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) {
            output[base + i] = items[i];
        }
    }
}

__device__ __forceinline__ void load_items(const Int4* input, int64_t base, int64_t n, bool, Int4 (&items)[kScanItems]) {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        items[i] = base + i < n ? input[base + i] : zero_value<Int4>();
    }
}

__device__ __forceinline__ void store_items(Int4* output, int64_t base, int64_t n, bool, const Int4 (&items)[kScanItems]) {
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) {
            output[base + i] = items[i];
        }
    }
}

// Exclusive scan of each kScanTile block, offset by the sum of all earlier blocks: with
// sum_totals the block adds up block_totals[0 .. block) itself, otherwise block_totals already
// holds the exclusive scan of the totals and entry [block] is the offset. In-place use (input == output) is allowed: every
// thread reads its items before writing them.
template<typename T>
__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(const T* input, T* output,
    const T* block_totals, int64_t n, int sum_totals, int vector) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    T items[kScanItems];
    load_items(input, base, n, vector != 0, items);
    T offset = zero_value<T>();
    if (block_totals && sum_totals) {
        T partial = zero_value<T>();
        for (int i = threadIdx.x; i < static_cast<int>(blockIdx.x); i += kScanThreads) {
            partial = add(partial, block_totals[i]);
        }
        offset = block_sum(partial);
        __syncthreads();  // the block scan below reuses the LDS of block_sum
    }
    else if (block_totals) {
        offset = block_totals[blockIdx.x];
    }
    T sum = zero_value<T>();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        sum = add(sum, items[i]);
    }
    T total;
    T running = add(offset, block_exclusive_scan(sum, &total));
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        T value = items[i];
        items[i] = running;
        running = add(running, value);
    }
    store_items(output, base, n, vector != 0, items);
}

// Up to this many blocks, each downsweep block sums the earlier block totals itself (at most
// kSumTotalsBlocks / kScanThreads loads per thread) instead of a separate scan of the totals.
constexpr int64_t kSumTotalsBlocks = 2048;

template<typename T>
int exclusive_scan(gsn_context* context, const T* input, T* output, int64_t n, int depth) {
    if (n <= 0) {
        return GSN_OK;
    }
    const int vector = (reinterpret_cast<uintptr_t>(input) % 16 == 0 && reinterpret_cast<uintptr_t>(output) % 16 == 0) ? 1 : 0;
    int64_t blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks == 1) {
        scan_downsweep_kernel<T><<<1, kScanThreads, 0, context->stream>>>(input, output, nullptr, n, 0, vector);
        return check_launch("scan_downsweep_kernel");
    }
    if (depth >= 4) {
        return fail(GSN_ERROR_CAPACITY, "exclusive_scan: input too large (%lld)", static_cast<long long>(n));
    }
    // Block totals live in a per-depth region of the scan scratch slot.
    size_t level_bytes = static_cast<size_t>(blocks) * sizeof(T);
    size_t offset_bytes = 0;
    int64_t level = n;
    for (int d = 0; d < depth; ++d) {
        level = (level + kScanTile - 1) / kScanTile;
        offset_bytes += ((static_cast<size_t>(level) * sizeof(T) + 255) & ~static_cast<size_t>(255));
    }
    size_t needed = offset_bytes + ((level_bytes + 255) & ~static_cast<size_t>(255)) + 4096 * sizeof(T);
    char* base = static_cast<char*>(scratch(context, kSlotScanBlocks, needed * 2 + (1 << 20)));
    if (!base) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    T* totals = reinterpret_cast<T*>(base + offset_bytes);
    scan_reduce_kernel<T><<<static_cast<unsigned>(blocks), kScanThreads, 0, context->stream>>>(input, totals, n);
    GSN_TRY(check_launch("scan_reduce_kernel"));
    const int sum_totals = blocks <= kSumTotalsBlocks ? 1 : 0;
    if (!sum_totals) {
        GSN_TRY(exclusive_scan<T>(context, totals, totals, blocks, depth + 1));
    }
    scan_downsweep_kernel<T><<<static_cast<unsigned>(blocks), kScanThreads, 0, context->stream>>>(input, output, totals,
        n, sum_totals, vector);
    return check_launch("scan_downsweep_kernel");
}

} // namespace

int exclusive_scan_int(gsn_context* context, const int* input, int* output, int64_t n) {
    return exclusive_scan<int>(context, input, output, n, 0);
}

int exclusive_scan_int_from_totals(gsn_context* context, const int* input, int* output, int64_t n,
    const int* block_totals) {
    const int64_t blocks = (n + kScanTile - 1) / kScanTile;
    if (n <= 0 || blocks == 1 || blocks > kSumTotalsBlocks) {
        return exclusive_scan<int>(context, input, output, n, 0);
    }
    const int vector = (reinterpret_cast<uintptr_t>(input) % 16 == 0 && reinterpret_cast<uintptr_t>(output) % 16 == 0) ? 1 : 0;
    scan_downsweep_kernel<int><<<static_cast<unsigned>(blocks), kScanThreads, 0, context->stream>>>(input, output,
        block_totals, n, 1, vector);
    return check_launch("scan_downsweep_kernel");
}

int exclusive_scan_int4(gsn_context* context, const Int4* input, Int4* output, int64_t n) {
    return exclusive_scan<Int4>(context, input, output, n, 0);
}

} // namespace gsn
