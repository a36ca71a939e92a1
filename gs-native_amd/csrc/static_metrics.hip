// Allocation-free image metrics: the same kernels as gsn_image_metrics, with the per-block partial
// sums in a file-scope device array instead of context scratch, and the results left in device
// memory (no host synchronisation). For callers that cannot allocate device memory: code under
// HIP graph capture, evaluation loops that must not call hipMalloc.
//
// The array is resolved once per device by gsn_context_create (hipGetSymbolAddress with that
// device current); launches use that address, never the array's host-side name. Every call on a
// device shares the device's copy of the array, so calls that overlap on one device race.
#include "gsn_internal.h"

#include <mutex>

// Per-block sums (3 doubles per 32x16-pixel block and channel) of the largest image accepted:
// 131,072 blocks = 3 MiB of device memory held on each device the library runs on, enough for
// about 22 million pixels (e.g. 5187 x 3361).
constexpr int64_t kStaticMetricBlocks = 131072;
static __device__ __align__(256) double scratch[kStaticMetricBlocks * 3];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_address[kMaxDevices] = {};

} // namespace

namespace gsn {

int resolve_static_scratch(int device, void** address) {
    *address = nullptr;
    if (device < 0 || device >= kMaxDevices) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "device %d is beyond the %d devices the static scratch supports",
            device, kMaxDevices);
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_address[device]) {
        void* resolved = nullptr;
        hipError_t error = hipGetSymbolAddress(&resolved, HIP_SYMBOL(::scratch));
        if (error != hipSuccess || !resolved) {
            (void)hipGetLastError();
            return fail(GSN_ERROR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                hipGetErrorString(error));
        }
        g_scratch_address[device] = resolved;
    }
    *address = g_scratch_address[device];
    return GSN_OK;
}

} // namespace gsn

extern "C" int gsn_image_metrics_static(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, double* device_out) {
    using namespace gsn;
    if (!context || width <= 0 || height <= 0 || !image || !target || !device_out) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_image_metrics_static: invalid arguments");
    }
    const int64_t blocks = image_metric_blocks(width, height);
    if (blocks > kStaticMetricBlocks) {
        return fail(GSN_ERROR_CAPACITY,
            "gsn_image_metrics_static: a %dx%d image needs %lld partial-sum blocks, above the static maximum of %lld "
            "(about 22 million pixels)",
            width, height, static_cast<long long>(blocks), static_cast<long long>(kStaticMetricBlocks));
    }
    if (!context->static_scratch) {
        return fail(GSN_ERROR_HIP, "gsn_image_metrics_static: the static scratch of device %d is not resolved",
            context->device);
    }
    return launch_image_metrics(context, width, height, image, target, static_cast<double*>(context->static_scratch),
        device_out);
}
