// Stable LSD radix sort of (uint32 key, uint32 value) pairs for wave64 / gfx950.
//
// Each pass (<= 8 digit bits) runs three kernels over 4096-element tiles:
//   1. radix_histogram: per-tile digit histogram in LDS, written bin-major [bin][tile].
//   2. radix_digit_scan: one workgroup per digit scans its row of the table in place (the
//      digit's start inside every tile's share) and writes the row total.
//   3. radix_scatter: adds the exclusive prefix of the row totals (the digit's global start),
//      then stable in-tile ranking with wave ballots (digit match masks), local reordering
//      through LDS, and a write-out in which consecutive lanes store consecutive positions of a
//      digit run.
// The row scan replaces a device-wide scan of the whole table (two launches per pass).
// Stability: a tile is split into eight 512-element wave chunks processed in index order, and
// ranks are offset by the digit counts of earlier waves, so equal digits keep input order.
#include "gsn_internal.h"

namespace gsn {
namespace {

// 512 threads x 8 items: eight waves rank a tile (measured against 256 x 16: depth sort 0.172 ->
// 0.162 ms, tile sort 0.292 -> 0.280 ms at 3M Gaussians / 30M records).
constexpr int kSortThreads = 512;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
static_assert(kSortTile == kRadixTile, "histogram columns cover kRadixTile elements");
constexpr int kWaves = kSortThreads / 64;
constexpr int kWaveChunk = kSortTile / kWaves;        // 512

__global__ __launch_bounds__(kSortThreads) void radix_histogram_kernel(const uint32_t* keys, int64_t n, int shift,
    int bins, int tiles, int* histogram) {
    __shared__ int counts[256];
    int t = threadIdx.x;
    if (t < 256) {
        counts[t] = 0;
    }
    __syncthreads();
    int64_t base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    uint32_t mask = static_cast<uint32_t>(bins - 1);
    uint32_t loaded[kSortItems];
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        int64_t index = base + i * kSortThreads + t;
        loaded[i] = 0u;
        if (index < n) {
            loaded[i] = keys[index];
        }
    }
#pragma unroll
    for (int i = 0; i < kSortItems; ++i) {
        if (base + i * kSortThreads + t < n) {
            atomicAdd(&counts[(loaded[i] >> shift) & mask], 1);
        }
    }
    __syncthreads();
    if (t < bins) {
        histogram[static_cast<int64_t>(t) * tiles + blockIdx.x] = counts[t];
    }
}

// Exclusive scan of digit blockIdx.x's row of the [bin][tile] table, in place, in chunks of
// kSortTile entries (eight consecutive entries per thread); the row total goes to totals.
__global__ __launch_bounds__(kSortThreads) void radix_digit_scan_kernel(int* histogram, int tiles, int* totals) {
    __shared__ int s_wave[kWaves];
    int* row = histogram + static_cast<int64_t>(blockIdx.x) * tiles;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    int carry = 0;
    for (int chunk = 0; chunk < tiles; chunk += kSortTile) {
        int values[kSortItems];
        int sum = 0;
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            int index = chunk + t * kSortItems + i;
            values[i] = index < tiles ? row[index] : 0;
            sum += values[i];
        }
        int inclusive = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            int other = __shfl_up(inclusive, d, 64);
            if (lane >= d) {
                inclusive += other;
            }
        }
        if (lane == 63) {
            s_wave[wave] = inclusive;
        }
        __syncthreads();
        int before = 0;
        int chunk_total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            int v = s_wave[w];
            before += w < wave ? v : 0;
            chunk_total += v;
        }
        int running = carry + before + inclusive - sum;
#pragma unroll
        for (int i = 0; i < kSortItems; ++i) {
            int index = chunk + t * kSortItems + i;
            if (index < tiles) {
                row[index] = running;
            }
            running += values[i];
        }
        carry += chunk_total;
        // s_wave is rewritten by the next chunk.
        __syncthreads();
    }
    if (t == 0) {
        totals[blockIdx.x] = carry;
    }
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint32_t* keys_in, const uint32_t* values_in,
    int64_t n, int shift, int bits, int tiles, const int* offsets, const int* totals, uint32_t* keys_out,
    uint32_t* values_out) {
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_values[kSortTile];
    __shared__ int s_wave_count[kWaves][256];
    __shared__ int s_digit_start[256];
    __shared__ int s_global[256];
    __shared__ int s_scan[256];
    __shared__ int s_base[256];

    const int t = threadIdx.x;
    const int wave = t >> 6;
    const int lane = t & 63;
    const int bins = 1 << bits;
    const uint32_t mask = static_cast<uint32_t>(bins - 1);
    const int64_t tile_base = static_cast<int64_t>(blockIdx.x) * kSortTile;
    const int valid = static_cast<int>(n - tile_base < kSortTile ? n - tile_base : kSortTile);
    const uint64_t lanemask_lt = (1ull << lane) - 1ull;

    for (int d = lane; d < 256; d += 64) {
        s_wave_count[wave][d] = 0;
    }
    __syncthreads();

    uint32_t key_reg[kSortItems];
    uint32_t value_reg[kSortItems];
    int rank_reg[kSortItems];
    // All loads first (16 per lane in flight), then the ranking, which serialises on LDS counters.
    // Each pass reads its input once: non-temporal loads (tile sort at 30M records 0.339 ->
    // 0.285 ms; non-temporal stores of the scattered output measured 0.52 ms and stay ordinary).
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        int local = wave * kWaveChunk + j * 64 + lane;
        key_reg[j] = 0xffffffffu;
        value_reg[j] = 0u;
        if (local < valid) {
            key_reg[j] = __builtin_nontemporal_load(keys_in + tile_base + local);
            value_reg[j] = __builtin_nontemporal_load(values_in + tile_base + local);
        }
    }
#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        int local = wave * kWaveChunk + j * 64 + lane;
        bool in_range = local < valid;
        uint32_t key = key_reg[j];
        // Out-of-range slots take the largest digit so that they rank after every valid element.
        uint32_t digit = in_range ? ((key >> shift) & mask) : mask;
        uint64_t peers = ~0ull;
        for (int b = 0; b < bits; ++b) {
            uint64_t vote = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? vote : ~vote;
        }
        int before = __popcll(peers & lanemask_lt);
        int count = s_wave_count[wave][digit];
        rank_reg[j] = count + before;
        if (before == 0) {
            s_wave_count[wave][digit] = count + __popcll(peers);
        }
        // Keeps the next round's counter reads after this round's leader writes.
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // Per digit: exclusive prefix across waves (in place) and the tile total; the digit's start
    // in this tile's share of the output (its row-scanned table entry) and its row total.
    int digit_total = 0;
    int row_total = 0;
    if (t < bins) {
        int running = 0;
        for (int w = 0; w < kWaves; ++w) {
            int c = s_wave_count[w][t];
            s_wave_count[w][t] = running;
            running += c;
        }
        digit_total = running;
        row_total = totals[t];
        s_global[t] = offsets[static_cast<int64_t>(t) * tiles + blockIdx.x];
    }
    if (t < 256) {
        s_scan[t] = digit_total;
        s_base[t] = row_total;
    }
    __syncthreads();
    // Inclusive scans over the digits of the in-tile counts and of the row totals.
    for (int offset = 1; offset < 256; offset <<= 1) {
        int other = (t < 256 && t >= offset) ? s_scan[t - offset] : 0;
        int other_base = (t < 256 && t >= offset) ? s_base[t - offset] : 0;
        __syncthreads();
        if (t < 256) {
            s_scan[t] += other;
            s_base[t] += other_base;
        }
        __syncthreads();
    }
    if (t < 256) {
        s_digit_start[t] = s_scan[t] - digit_total;
    }
    if (t < bins) {
        s_global[t] += s_base[t] - row_total;
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kSortItems; ++j) {
        int local = wave * kWaveChunk + j * 64 + lane;
        uint32_t digit = local < valid ? ((key_reg[j] >> shift) & mask) : mask;
        int position = s_digit_start[digit] + s_wave_count[wave][digit] + rank_reg[j];
        s_keys[position] = key_reg[j];
        s_values[position] = value_reg[j];
    }
    __syncthreads();

    for (int i = t; i < valid; i += kSortThreads) {
        uint32_t key = s_keys[i];
        uint32_t digit = (key >> shift) & mask;
        int64_t destination = static_cast<int64_t>(s_global[digit]) + (i - s_digit_start[digit]);
        if (keys_out) {
            keys_out[destination] = key;
        }
        values_out[destination] = s_values[i];
    }
}

} // namespace

// The [256][tiles] table, followed by 256 row totals.
int* radix_histogram_buffer(gsn_context* context, int64_t n) {
    int64_t tiles = (n + kSortTile - 1) / kSortTile;
    return static_cast<int*>(
        scratch(context, kSlotSortHistogram, (static_cast<size_t>(tiles) + 1) * 256 * sizeof(int)));
}

// Digit widths: the narrower digits come first (13 tile bits: 6 then 7). The first pass of the
// tile sort reads the records in emission order, where neighbouring records share tiles; fewer
// bins there measured slightly faster (tile sort 0.298 -> 0.291 ms at 30M records).
int radix_first_digit_bits(int bits) {
    bits = bits < 1 ? 1 : (bits > 32 ? 32 : bits);
    int passes = (bits + 7) / 8;
    return bits / passes;
}

int radix_sort_pairs(gsn_context* context, uint32_t* keys, uint32_t* values, int64_t n, int bits,
    uint32_t* keys_out, uint32_t* values_out, bool first_histogram_ready) {
    if (n <= 0) {
        return GSN_OK;
    }
    if (n > 0x7fffffffll) {
        return fail(GSN_ERROR_CAPACITY, "radix_sort_pairs: %lld elements exceed int32 positions", static_cast<long long>(n));
    }
    if (bits < 1) {
        bits = 1;
    }
    if (bits > 32) {
        bits = 32;
    }
    int passes = (bits + 7) / 8;
    int tiles = static_cast<int>((n + kSortTile - 1) / kSortTile);
    size_t elements = static_cast<size_t>(n);
    uint32_t* scratch_keys = static_cast<uint32_t*>(scratch(context, kSlotSortKeysA, elements * sizeof(uint32_t)));
    uint32_t* scratch_values = static_cast<uint32_t*>(scratch(context, kSlotSortValuesA, elements * sizeof(uint32_t)));
    int* histogram = radix_histogram_buffer(context, n);
    if (!scratch_keys || !scratch_values || !histogram) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    int* totals = histogram + static_cast<int64_t>(tiles) * 256;
    uint32_t* final_keys = keys_out;
    if (!final_keys) {
        final_keys = static_cast<uint32_t*>(scratch(context, kSlotSortKeysB, elements * sizeof(uint32_t)));
        if (!final_keys) {
            return GSN_ERROR_OUT_OF_MEMORY;
        }
    }
    // Pass p reads src and writes dst; the last pass writes the caller's outputs (which must not
    // alias the inputs), earlier passes alternate scratch, input, scratch, ... so that no pass
    // reads and writes the same buffer.
    uint32_t* src_keys = keys;
    uint32_t* src_values = values;
    int shift = 0;
    for (int pass = 0; pass < passes; ++pass) {
        int remaining = bits - shift;
        int digit_bits = remaining / (passes - pass);
        int bins = 1 << digit_bits;
        bool last = pass == passes - 1;
        bool to_scratch = (pass % 2) == 0;
        uint32_t* dst_keys = last ? final_keys : (to_scratch ? scratch_keys : keys);
        uint32_t* dst_values = last ? values_out : (to_scratch ? scratch_values : values);
        if (pass > 0 || !first_histogram_ready) {
            radix_histogram_kernel<<<tiles, kSortThreads, 0, context->stream>>>(src_keys, n, shift, bins, tiles,
                histogram);
            GSN_TRY(check_launch("radix_histogram_kernel"));
        }
        radix_digit_scan_kernel<<<bins, kSortThreads, 0, context->stream>>>(histogram, tiles, totals);
        GSN_TRY(check_launch("radix_digit_scan_kernel"));
        radix_scatter_kernel<<<tiles, kSortThreads, 0, context->stream>>>(src_keys, src_values, n, shift, digit_bits, tiles,
            histogram, totals, dst_keys, dst_values);
        GSN_TRY(check_launch("radix_scatter_kernel"));
        src_keys = dst_keys;
        src_values = dst_values;
        shift += digit_bits;
    }
    return GSN_OK;
}

} // namespace gsn

extern "C" int gsn_selftest_sort_pairs(gsn_context* context, uint32_t* keys, uint32_t* values, int count, int bits,
    uint32_t* keys_out, uint32_t* values_out) {
    using namespace gsn;
    if (!context || count < 0 || bits < 1 || bits > 32 ||
        (count > 0 && (!keys || !values || !keys_out || !values_out || keys_out == keys || values_out == values))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_selftest_sort_pairs: invalid arguments");
    }
    return radix_sort_pairs(context, keys, values, count, bits, keys_out, values_out);
}
