// Device-memory helpers of the host classes (HIP runtime host API).
#pragma once

#include "gsn_hip.h"

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace gs_native {

bool DeviceAllocate(void** pointer, size_t bytes);
void DeviceFree(void* pointer);
// Logs and returns false for a non-zero status of a C-ABI call.
bool Check(int status, const char* what);

template<typename T>
bool DeviceAllocateCount(T** pointer, size_t count) {
    void* raw = nullptr;
    bool ok = DeviceAllocate(&raw, count * sizeof(T));
    *pointer = static_cast<T*>(raw);
    return ok;
}

inline const gsn_camera* AsCamera(const void* camera_params) {
    return static_cast<const gsn_camera*>(camera_params);
}

} // namespace gs_native
