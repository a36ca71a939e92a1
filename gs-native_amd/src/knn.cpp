// Host entry of the 3-NN initialisation (reference src/core/knn/knn.cu:180-255).
#include "core/knn/knn.h"

#include "device_memory.h"
#include "log.h"
#include "util/device.h"

namespace gs_native {

std::vector<float> ComputeMeanDist2Flat(const float* xyz, size_t count) {
    if (count < 2) {
        return std::vector<float>(count, 0.0f);
    }
    if (count > static_cast<size_t>(0x7fffffff)) {
        GS_NATIVE_ERROR("KNN failed: %zu points exceed the int range", count);
        return {};
    }
    gsn_context* context = DefaultContext();
    float* points = nullptr;
    float* distances = nullptr;
    std::vector<float> result(count, 0.0f);
    bool ok = context && DeviceAllocateCount(&points, 3 * count) && DeviceAllocateCount(&distances, count) &&
        Check(gsn_memcpy_to_device(points, xyz, 3 * count * sizeof(float)), "gsn_memcpy_to_device") &&
        Check(gsn_knn3(context, points, static_cast<int>(count), nullptr, distances), "gsn_knn3") &&
        Check(gsn_context_synchronize(context), "gsn_context_synchronize") &&
        Check(gsn_memcpy_to_host(result.data(), distances, count * sizeof(float)), "gsn_memcpy_to_host");
    DeviceFree(points);
    DeviceFree(distances);
    if (!ok) {
        GS_NATIVE_ERROR("KNN failed");
        return {};
    }
    return result;
}

} // namespace gs_native
