// Device-wide exclusive prefix scan (reduce -> scan of block totals -> downsweep).
// Used for record offsets (int[N+1]), density decisions (int4[N+1]) and radix-sort histograms.
#include "gsn_internal.h"

namespace gsn {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

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

// Exclusive scan of one value per thread across the block; returns the block total via *total.
template<typename T>
__device__ T block_exclusive_scan(T value, T* total) {
    __shared__ T lds[kScanThreads];
    int t = threadIdx.x;
    lds[t] = value;
    __syncthreads();
    // Hillis-Steele inclusive scan in LDS (8 steps for 256 threads).
    for (int offset = 1; offset < kScanThreads; offset <<= 1) {
        T other = t >= offset ? lds[t - offset] : zero_value<T>();
        __syncthreads();
        if (t >= offset) {
            lds[t] = add(lds[t], other);
        }
        __syncthreads();
    }
    T inclusive = lds[t];
    *total = lds[kScanThreads - 1];
    __syncthreads();
    T exclusive = t == 0 ? zero_value<T>() : lds[t - 1];
    __syncthreads();
    (void)inclusive;
    return exclusive;
}

template<typename T>
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const T* input, T* block_totals, int64_t n) {
    int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    T sum = zero_value<T>();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t index = base + i;
        if (index < n) {
            sum = add(sum, input[index]);
        }
    }
    T total;
    block_exclusive_scan(sum, &total);
    if (threadIdx.x == 0) {
        block_totals[blockIdx.x] = total;
    }
}

template<typename T>
__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(const T* input, T* output,
    const T* block_offsets, int64_t n) {
    int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    T items[kScanItems];
    T sum = zero_value<T>();
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t index = base + i;
        items[i] = index < n ? input[index] : zero_value<T>();
        sum = add(sum, items[i]);
    }
    T total;
    T running = block_exclusive_scan(sum, &total);
    if (block_offsets) {
        running = add(running, block_offsets[blockIdx.x]);
    }
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        int64_t index = base + i;
        if (index < n) {
            output[index] = running;
        }
        running = add(running, items[i]);
    }
}

template<typename T>
int exclusive_scan(gsn_context* context, const T* input, T* output, int64_t n, int depth) {
    if (n <= 0) {
        return GSN_OK;
    }
    int64_t blocks = (n + kScanTile - 1) / kScanTile;
    if (blocks == 1) {
        scan_downsweep_kernel<T><<<1, kScanThreads, 0, context->stream>>>(input, output, nullptr, n);
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
    GSN_TRY(exclusive_scan<T>(context, totals, totals, blocks, depth + 1));
    scan_downsweep_kernel<T><<<static_cast<unsigned>(blocks), kScanThreads, 0, context->stream>>>(input, output, totals, n);
    return check_launch("scan_downsweep_kernel");
}

} // namespace

int exclusive_scan_int(gsn_context* context, const int* input, int* output, int64_t n) {
    return exclusive_scan<int>(context, input, output, n, 0);
}

int exclusive_scan_int4(gsn_context* context, const Int4* input, Int4* output, int64_t n) {
    return exclusive_scan<Int4>(context, input, output, n, 0);
}

} // namespace gsn
