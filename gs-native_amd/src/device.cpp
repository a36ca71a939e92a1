// Shared execution context of the gs_native classes and small device-memory helpers.
#include "device_memory.h"
#include "log.h"
#include "util/device.h"

#include <mutex>

namespace gs_native {

gsn_context* DefaultContext() {
    static gsn_context* context = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        if (gsn_context_create(&context) != GSN_OK) {
            GS_NATIVE_ERROR("creating the device context failed: %s", gsn_last_error());
            context = nullptr;
        }
    });
    return context;
}

bool DeviceAllocate(void** pointer, size_t bytes) {
    *pointer = nullptr;
    if (bytes == 0) {
        return true;
    }
    if (hipMalloc(pointer, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *pointer = nullptr;
        return false;
    }
    return true;
}

void DeviceFree(void* pointer) {
    if (pointer) {
        (void)hipFree(pointer);
    }
}

bool Check(int status, const char* what) {
    if (status != GSN_OK) {
        GS_NATIVE_ERROR("%s: %s (%s)", what, gsn_status_string(status), gsn_last_error());
        return false;
    }
    return true;
}

} // namespace gs_native
