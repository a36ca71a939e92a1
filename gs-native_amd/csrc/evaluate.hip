// Held-out evaluation of a render (reference demo/render.cpp:107-155): the float image is clamped
// to [0, 1], scaled by 255 in FP32 and rounded half to even to 8 bits (OpenCV's convertTo to
// CV_8UC3 with scale 255), and compared with the 8-bit target by the sum of squared differences
// over all pixels and channels; PSNR = -10 log10(SSE / (3 W H 255^2)), infinite for SSE = 0.
// Integer work: the sum is exact, so the result does not depend on the summation order.
#include "gsn_internal.h"

#include <cmath>

namespace gsn {
namespace {

constexpr int kEvalThreads = 256;

__global__ __launch_bounds__(kEvalThreads) void quantized_error_kernel(const float* __restrict__ image,
    const uint8_t* __restrict__ target, uint8_t* __restrict__ quantized, int64_t values,
    unsigned long long* __restrict__ sse) {
    __shared__ unsigned long long partial[kEvalThreads / 64];
    unsigned long long sum = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kEvalThreads + threadIdx.x; i < values;
         i += static_cast<int64_t>(gridDim.x) * kEvalThreads) {
        float v = fminf(fmaxf(image[i], 0.0f), 1.0f);
        int q = static_cast<int>(rintf(v * 255.0f));
        if (quantized) {
            quantized[i] = static_cast<uint8_t>(q);
        }
        int d = q - static_cast<int>(target[i]);
        sum += static_cast<unsigned long long>(d * d);
    }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        sum += __shfl_xor(sum, offset);
    }
    if ((threadIdx.x & 63) == 0) {
        partial[threadIdx.x >> 6] = sum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long block = 0;
        for (int w = 0; w < kEvalThreads / 64; ++w) {
            block += partial[w];
        }
        atomicAdd(sse, block);
    }
}

} // namespace
} // namespace gsn

extern "C" int gsn_image_quantized_error(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, uint8_t* quantized, double* out) {
    using namespace gsn;
    if (!context || width <= 0 || height <= 0 || !image || !target || !out) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_image_quantized_error: invalid arguments");
    }
    const int64_t values = static_cast<int64_t>(width) * height * 3;
    auto* sse = static_cast<unsigned long long*>(scratch(context, kSlotQuantizedError, sizeof(unsigned long long)));
    if (!sse) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    GSN_HIP(hipMemsetAsync(sse, 0, sizeof(unsigned long long), context->stream));
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>(grid_for(values, kEvalThreads), 8192));
    quantized_error_kernel<<<blocks, kEvalThreads, 0, context->stream>>>(image, target, quantized, values, sse);
    GSN_TRY(check_launch("quantized_error_kernel"));
    auto* host = reinterpret_cast<unsigned long long*>(static_cast<char*>(context->pinned) + 64);
    GSN_HIP(hipMemcpyAsync(host, sse, sizeof(unsigned long long), hipMemcpyDeviceToHost, context->stream));
    GSN_HIP(hipStreamSynchronize(context->stream));
    const double total = static_cast<double>(*host);
    out[0] = total;
    out[1] = total == 0.0 ? INFINITY : -10.0 * std::log10(total / (3.0 * width * height * 255.0 * 255.0));
    return GSN_OK;
}
