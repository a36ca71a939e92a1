// Context lifetime, status reporting, scratch management and device-memory plumbing.
#include "gsn_internal.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

namespace {

thread_local char g_last_error[512] = "";

} // namespace

namespace gsn {

int fail(int status, const char* format, ...) {
    va_list args;
    va_start(args, format);
    vsnprintf(g_last_error, sizeof(g_last_error), format, args);
    va_end(args);
    return status;
}

int check_hip(hipError_t error, const char* what) {
    if (error == hipSuccess) {
        return GSN_OK;
    }
    int status = error == hipErrorOutOfMemory ? GSN_ERROR_OUT_OF_MEMORY : GSN_ERROR_HIP;
    return fail(status, "%s: %s", what, hipGetErrorString(error));
}

int check_launch(const char* kernel) {
    hipError_t error = hipGetLastError();
    if (error == hipSuccess) {
        return GSN_OK;
    }
    return fail(GSN_ERROR_HIP, "launch of %s failed: %s", kernel, hipGetErrorString(error));
}

void* scratch(gsn_context* context, ScratchSlot slot, size_t bytes) {
    if (bytes == 0) {
        bytes = 256;
    }
    if (context->slot_bytes[slot] >= bytes) {
        return context->slots[slot];
    }
    // Grow by 25% headroom so that slowly growing record counts do not reallocate every call.
    size_t grown = bytes + bytes / 4;
    grown = (grown + 255) & ~static_cast<size_t>(255);
    if (context->slots[slot]) {
        (void)hipStreamSynchronize(context->stream);
        if (context->aux_stream) {
            (void)hipStreamSynchronize(context->aux_stream);
        }
        (void)hipFree(context->slots[slot]);
        context->slots[slot] = nullptr;
        context->slot_bytes[slot] = 0;
    }
    void* pointer = nullptr;
    if (hipMalloc(&pointer, grown) != hipSuccess) {
        (void)hipGetLastError();
        fail(GSN_ERROR_OUT_OF_MEMORY, "scratch slot %d: cannot allocate %zu bytes", static_cast<int>(slot), grown);
        return nullptr;
    }
    context->slots[slot] = pointer;
    context->slot_bytes[slot] = grown;
    return pointer;
}

} // namespace gsn

extern "C" {

int gsn_context_create(gsn_context** context) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_create: null output");
    }
    *context = nullptr;
    gsn_context* created = new (std::nothrow) gsn_context();
    if (!created) {
        return gsn::fail(GSN_ERROR_OUT_OF_MEMORY, "gsn_context_create: host allocation failed");
    }
    int status = gsn::check_hip(hipGetDevice(&created->device), "hipGetDevice");
    if (status == GSN_OK) {
        status = gsn::check_hip(hipDeviceGetAttribute(&created->cu_count, hipDeviceAttributeMultiprocessorCount,
                                    created->device),
            "hipDeviceGetAttribute");
    }
    if (status == GSN_OK) {
        status = gsn::resolve_static_scratch(created->device, &created->static_scratch);
    }
    if (status == GSN_OK) {
        status = gsn::check_hip(hipStreamCreate(&created->stream), "hipStreamCreate");
    }
    if (status == GSN_OK) {
        status = gsn::check_hip(hipHostMalloc(&created->pinned, 256, hipHostMallocDefault), "hipHostMalloc");
    }
    if (status == GSN_OK) {
        status = gsn::check_hip(hipEventCreateWithFlags(&created->readback_event, hipEventDisableTiming),
            "hipEventCreateWithFlags");
    }
    if (status != GSN_OK) {
        gsn_context_destroy(created);
        return status;
    }
    *context = created;
    return GSN_OK;
}

void gsn_context_destroy(gsn_context* context) {
    if (!context) {
        return;
    }
    if (context->stream) {
        (void)hipStreamSynchronize(context->stream);
    }
    if (context->aux_stream) {
        (void)hipStreamSynchronize(context->aux_stream);
        (void)hipStreamDestroy(context->aux_stream);
    }
    if (context->fork_event) {
        (void)hipEventDestroy(context->fork_event);
    }
    if (context->join_event) {
        (void)hipEventDestroy(context->join_event);
    }
    for (int slot = 0; slot < kSlotCount; ++slot) {
        if (context->slots[slot]) {
            (void)hipFree(context->slots[slot]);
        }
    }
    if (context->pinned) {
        (void)hipHostFree(context->pinned);
    }
    if (context->readback_event) {
        (void)hipEventDestroy(context->readback_event);
    }
    if (context->raster_pairs) {
        (void)hipFree(context->raster_pairs);
    }
    for (int k = 0; k < 2 * context->event_capacity; ++k) {
        if (context->events[k]) {
            (void)hipEventDestroy(context->events[k]);
        }
    }
    delete[] context->events;
    delete[] context->event_stage;
    if (context->stream) {
        (void)hipStreamDestroy(context->stream);
    }
    delete context;
}

int gsn_device_synchronize(void) {
    return gsn::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int gsn_context_synchronize(gsn_context* context) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_synchronize: null context");
    }
    if (context->aux_stream) {
        GSN_HIP(hipStreamSynchronize(context->aux_stream));
    }
    return gsn::check_hip(hipStreamSynchronize(context->stream), "hipStreamSynchronize");
}

void* gsn_context_stream(gsn_context* context) {
    return context ? static_cast<void*>(context->stream) : nullptr;
}

size_t gsn_context_scratch_bytes(gsn_context* context) {
    size_t total = 0;
    if (context) {
        for (int slot = 0; slot < kSlotCount; ++slot) {
            total += context->slot_bytes[slot];
        }
    }
    return total;
}

const char* gsn_status_string(int status) {
    switch (status) {
    case GSN_OK:
        return "ok";
    case GSN_ERROR_INVALID_ARGUMENT:
        return "invalid argument";
    case GSN_ERROR_HIP:
        return "HIP runtime error";
    case GSN_ERROR_OUT_OF_MEMORY:
        return "out of device memory";
    case GSN_ERROR_CAPACITY:
        return "capacity exceeded";
    default:
        return "unknown status";
    }
}

const char* gsn_last_error(void) {
    return g_last_error;
}

int gsn_device_count(int* count) {
    if (!count) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_device_count: null output");
    }
    *count = 0;
    hipError_t error = hipGetDeviceCount(count);
    if (error != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
    }
    return GSN_OK;
}

int gsn_device_name(char* name, size_t capacity) {
    if (!name || capacity == 0) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_device_name: null output");
    }
    int device = 0;
    GSN_HIP(hipGetDevice(&device));
    hipDeviceProp_t properties;
    GSN_HIP(hipGetDeviceProperties(&properties, device));
    snprintf(name, capacity, "%s (%s, %d CUs)", properties.name, properties.gcnArchName,
        properties.multiProcessorCount);
    return GSN_OK;
}

int gsn_device_malloc(void** pointer, size_t bytes) {
    if (!pointer) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_device_malloc: null output");
    }
    *pointer = nullptr;
    if (bytes == 0) {
        return GSN_OK;
    }
    return gsn::check_hip(hipMalloc(pointer, bytes), "hipMalloc");
}

int gsn_device_free(void* pointer) {
    if (!pointer) {
        return GSN_OK;
    }
    // Work already enqueued on any stream may still read the allocation.
    GSN_HIP(hipDeviceSynchronize());
    return gsn::check_hip(hipFree(pointer), "hipFree");
}

int gsn_memcpy_to_device(void* destination, const void* source, size_t bytes) {
    if (bytes == 0) {
        return GSN_OK;
    }
    return gsn::check_hip(hipMemcpy(destination, source, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}

int gsn_memcpy_to_host(void* destination, const void* source, size_t bytes) {
    if (bytes == 0) {
        return GSN_OK;
    }
    return gsn::check_hip(hipMemcpy(destination, source, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}

int gsn_memcpy_device(void* destination, const void* source, size_t bytes) {
    if (bytes == 0) {
        return GSN_OK;
    }
    return gsn::check_hip(hipMemcpy(destination, source, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}

int gsn_memset_device(void* destination, int value, size_t bytes) {
    if (bytes == 0) {
        return GSN_OK;
    }
    return gsn::check_hip(hipMemset(destination, value, bytes), "hipMemset");
}

int gsn_device_memory_info(size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_device_memory_info: null output");
    }
    return gsn::check_hip(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int gsn_event_create(void** event) {
    if (!event) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_event_create: null output");
    }
    hipEvent_t created = nullptr;
    GSN_HIP(hipEventCreate(&created));
    *event = created;
    return GSN_OK;
}

int gsn_event_destroy(void* event) {
    if (!event) {
        return GSN_OK;
    }
    return gsn::check_hip(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
}

int gsn_event_record(gsn_context* context, void* event) {
    if (!context || !event) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_event_record: null argument");
    }
    return gsn::check_hip(hipEventRecord(static_cast<hipEvent_t>(event), context->stream), "hipEventRecord");
}

int gsn_event_elapsed_ms(void* start, void* end, float* milliseconds) {
    if (!start || !end || !milliseconds) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_event_elapsed_ms: null argument");
    }
    GSN_HIP(hipEventSynchronize(static_cast<hipEvent_t>(end)));
    return gsn::check_hip(hipEventElapsedTime(milliseconds, static_cast<hipEvent_t>(start),
                              static_cast<hipEvent_t>(end)),
        "hipEventElapsedTime");
}

} // extern "C"

namespace gsn {

StageScope::StageScope(gsn_context* context, Stage stage, bool kernel_events)
    : context_(context), pair_(-1), kernel_events_(kernel_events) {
    if (!context || !((context->timing_mask >> stage) & 1u)) {
        return;
    }
    if (context->event_used == context->event_capacity) {
        int capacity = context->event_capacity ? 2 * context->event_capacity : 256;
        hipEvent_t* events = new (std::nothrow) hipEvent_t[2 * capacity];
        int* stages = new (std::nothrow) int[capacity];
        if (!events || !stages) {
            delete[] events;
            delete[] stages;
            return;
        }
        for (int k = 0; k < 2 * context->event_capacity; ++k) {
            events[k] = context->events[k];
        }
        for (int k = 0; k < context->event_capacity; ++k) {
            stages[k] = context->event_stage[k];
        }
        for (int k = 2 * context->event_capacity; k < 2 * capacity; ++k) {
            if (hipEventCreate(&events[k]) != hipSuccess) {
                events[k] = nullptr;
            }
        }
        delete[] context->events;
        delete[] context->event_stage;
        context->events = events;
        context->event_stage = stages;
        context->event_capacity = capacity;
    }
    pair_ = context->event_used++;
    context->event_stage[pair_] = stage;
    if (!kernel_events_ && context->events[2 * pair_]) {
        (void)hipEventRecord(context->events[2 * pair_], context->stream);
    }
}

StageScope::~StageScope() {
    if (pair_ >= 0 && !kernel_events_ && context_->events[2 * pair_ + 1]) {
        (void)hipEventRecord(context_->events[2 * pair_ + 1], context_->stream);
    }
}

hipEvent_t StageScope::start_event() const {
    return pair_ >= 0 ? context_->events[2 * pair_] : nullptr;
}

hipEvent_t StageScope::stop_event() const {
    return pair_ >= 0 ? context_->events[2 * pair_ + 1] : nullptr;
}

} // namespace gsn

extern "C" int gsn_context_set_timing(gsn_context* context, int enabled) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_set_timing: null context");
    }
    context->timing_mask = enabled != 0 ? (1u << kStageCount) - 1u : 0u;
    context->event_used = 0;
    return GSN_OK;
}

extern "C" int gsn_context_set_timing_mask(gsn_context* context, unsigned stage_mask) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_set_timing_mask: null context");
    }
    context->timing_mask = stage_mask & ((1u << kStageCount) - 1u);
    context->event_used = 0;
    return GSN_OK;
}

extern "C" int gsn_context_stage_times(gsn_context* context, double* milliseconds, int* launches, int stages) {
    if (!context || !milliseconds || !launches || stages < 0) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_stage_times: invalid arguments");
    }
    for (int s = 0; s < stages; ++s) {
        milliseconds[s] = 0.0;
        launches[s] = 0;
    }
    GSN_HIP(hipStreamSynchronize(context->stream));
    for (int pair = 0; pair < context->event_used; ++pair) {
        int stage = context->event_stage[pair];
        if (stage >= stages || !context->events[2 * pair] || !context->events[2 * pair + 1]) {
            continue;
        }
        float elapsed = 0.0f;
        if (hipEventElapsedTime(&elapsed, context->events[2 * pair], context->events[2 * pair + 1]) == hipSuccess) {
            milliseconds[stage] += elapsed;
            launches[stage] += 1;
        }
        else {
            (void)hipGetLastError();
        }
    }
    context->event_used = 0;
    return GSN_OK;
}

extern "C" int gsn_context_count_raster_pairs(gsn_context* context, int enabled) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_count_raster_pairs: null context");
    }
    if (enabled && !context->raster_pairs) {
        void* counter = nullptr;
        GSN_HIP(hipMalloc(&counter, sizeof(unsigned long long)));
        context->raster_pairs = static_cast<unsigned long long*>(counter);
        GSN_HIP(hipMemsetAsync(context->raster_pairs, 0, sizeof(unsigned long long), context->stream));
    }
    context->count_raster_pairs = enabled != 0;
    return GSN_OK;
}

extern "C" int gsn_context_prefetch_backward(gsn_context* context) {
    if (!context) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_prefetch_backward: null context");
    }
    if (!context->aux_stream) {
        // Default priority, like the main stream.
        GSN_HIP(hipStreamCreate(&context->aux_stream));
        GSN_HIP(hipEventCreateWithFlags(&context->fork_event, hipEventDisableTiming));
        GSN_HIP(hipEventCreateWithFlags(&context->join_event, hipEventDisableTiming));
    }
    context->prefetch_armed = true;
    return GSN_OK;
}

extern "C" int gsn_context_prefetch_reused(gsn_context* context, unsigned long long* calls) {
    if (!context || !calls) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_prefetch_reused: invalid arguments");
    }
    *calls = context->prefetch_reused;
    return GSN_OK;
}

extern "C" int gsn_context_raster_pairs(gsn_context* context, unsigned long long* pairs) {
    if (!context || !pairs) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_context_raster_pairs: invalid arguments");
    }
    *pairs = 0;
    if (!context->raster_pairs) {
        return GSN_OK;
    }
    // Bytes 128.. of the pinned buffer: the record-count and density readbacks use its start.
    void* host = static_cast<char*>(context->pinned) + 128;
    GSN_HIP(hipMemcpyAsync(host, context->raster_pairs, sizeof(unsigned long long), hipMemcpyDeviceToHost,
        context->stream));
    GSN_HIP(hipMemsetAsync(context->raster_pairs, 0, sizeof(unsigned long long), context->stream));
    GSN_HIP(hipStreamSynchronize(context->stream));
    memcpy(pairs, host, sizeof(unsigned long long));
    return GSN_OK;
}
