// Minimal logger with the reference's macro names (reference src/util/log.h:71-73). INFO lines
// are printed only when GS_NATIVE_LOG=info is set in the environment.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace gs_native {

inline bool InfoEnabled() {
    static const bool enabled = [] {
        const char* level = std::getenv("GS_NATIVE_LOG");
        return level && std::strcmp(level, "info") == 0;
    }();
    return enabled;
}

} // namespace gs_native

#define GS_NATIVE_INFO(...)                              \
    do {                                                 \
        if (::gs_native::InfoEnabled()) {                \
            std::fprintf(stderr, "[gs_native] " __VA_ARGS__); \
            std::fputc('\n', stderr);                    \
        }                                                \
    } while (0)
#define GS_NATIVE_WARN(...)                                       \
    do {                                                          \
        std::fprintf(stderr, "[gs_native] warning: " __VA_ARGS__); \
        std::fputc('\n', stderr);                                 \
    } while (0)
#define GS_NATIVE_ERROR(...)                                    \
    do {                                                        \
        std::fprintf(stderr, "[gs_native] error: " __VA_ARGS__); \
        std::fputc('\n', stderr);                               \
    } while (0)
