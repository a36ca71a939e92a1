// Shared device/host definitions for the gfx950 Gaussian-splatting training path:
// the SoA parameter layout, the camera record, renderer constants, and FP32 transcendental
// functions whose results are identical on host and device.
//
// Integer outputs of the path (tile rectangles, sort keys, density masks) depend on FP32
// exp/log/sin/cos. Vendor libm and the device math library round those differently in the last
// place, so this file supplies one implementation of each, built only from IEEE add/sub/mul,
// correctly rounded division/sqrt and rint, with FP contraction disabled inside each function:
//   * exp:  Cody-Waite reduction x = k*ln2 + r (ln2 split into a 13-bit head and a tail),
//           degree-7 Taylor polynomial of e^r on |r| <= 0.35, scaling by 2^k in two steps.
//   * log:  x = m*2^e with m in [sqrt(1/2), sqrt(2)), log(m) = 2*atanh(s), s = (m-1)/(m+1),
//           odd series to s^9.
//   * sin/cos: Cody-Waite reduction by pi/2 (three-part split, 13-bit heads), Taylor
//           polynomials on |r| <= pi/4, quadrant selection.
// Accuracy against glibc is checked in the oracle's tests (<= 2 ulp over the ranges used).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GSN_HD __host__ __device__ __forceinline__

namespace gsn {

// Parameter rows of the [58][N] SoA buffer (reference src/core/types/buffers.h:10-15).
constexpr int kPositionIndex = 0;
constexpr int kLogScaleIndex = 3;
constexpr int kRotationIndex = 6;
constexpr int kOpacityIndex = 9;
constexpr int kShIndex = 10;
constexpr int kParamCount = 58;
constexpr int kShCoeffCount = 16;

// Adam moments are the optimiser's own state, not part of the reference's interface, and are
// stored blocked: block b holds Gaussians [64 b, 64 b + 64), and inside it each of the 58 rows
// holds 64 (first, second) pairs. A wave of 64 consecutive Gaussians then reads and writes one
// contiguous, line-aligned 512-byte piece per row whatever the Gaussian count, and one row of
// both moments is one stream instead of two.
constexpr int kMomentBlock = 64;
constexpr int64_t kMomentBlockFloats = static_cast<int64_t>(kParamCount) * kMomentBlock * 2;

// Floats of the blocked moment buffer of `count` Gaussians.
GSN_HD int64_t moment_floats(int64_t count) {
    return (count + kMomentBlock - 1) / kMomentBlock * kMomentBlockFloats;
}

// Float index of the first moment of Gaussian i in row `row` (the second moment follows it).
GSN_HD int64_t moment_index(int64_t i, int row) {
    return (i / kMomentBlock) * kMomentBlockFloats + static_cast<int64_t>(row) * (kMomentBlock * 2) +
           (i % kMomentBlock) * 2;
}

// Renderer constants (reference src/core/render/renderer.cu:11-15).
constexpr int kTileSize = 16;
constexpr float kLowPassVariance = 0.3f;
constexpr float kFovClampScale = 1.3f;
constexpr float kMinNearDepth = 1e-6f;

// Real spherical-harmonic basis constants, degrees 0-3.
constexpr float kShC0 = 0.28209479177387814f;
constexpr float kShC1 = 0.4886025119029199f;
constexpr float kShC2_0 = 1.0925484305920792f;
constexpr float kShC2_1 = -1.0925484305920792f;
constexpr float kShC2_2 = 0.31539156525252005f;
constexpr float kShC2_3 = -1.0925484305920792f;
constexpr float kShC2_4 = 0.5462742152960396f;
constexpr float kShC3_0 = -0.5900435899266435f;
constexpr float kShC3_1 = 2.890611442640554f;
constexpr float kShC3_2 = -0.4570457994644658f;
constexpr float kShC3_3 = 0.3731763325901154f;
constexpr float kShC3_4 = -0.4570457994644658f;
constexpr float kShC3_5 = 1.445305721320277f;
constexpr float kShC3_6 = -0.5900435899266435f;

// Camera record; byte layout identical to gs_native::CameraParams (buffers.h:17-28).
struct Camera {
    float fx, fy, cx, cy;
    float R_c_w[9];
    float t_c_w[3];
    float near_depth, far_depth;
    int width, height;
    int64_t pixel_offset;
    int tile_count_x, tile_count_y;
    int tile_offset;
};
static_assert(sizeof(Camera) == 104, "Camera layout must match CameraParams");

GSN_HD float bits_to_float(uint32_t b) {
    union { uint32_t u; float f; } v;
    v.u = b;
    return v.f;
}

GSN_HD uint32_t float_to_bits(float f) {
    union { float f; uint32_t u; } v;
    v.f = f;
    return v.u;
}

// 2^n for n in [-126, 127].
GSN_HD float pow2i(int n) {
    return bits_to_float(static_cast<uint32_t>(n + 127) << 23);
}

GSN_HD float det_expf(float x) {
#pragma clang fp contract(off)
    if (x != x) {
        return x;
    }
    if (x > 88.72283935546875f) {
        return bits_to_float(0x7f800000u);
    }
    if (x < -103.97208404541015625f) {
        return 0.0f;
    }
    const float log2e = 0x1.715476p+0f;
    const float ln2_hi = 0x1.62ep-1f;
    const float ln2_lo = 0x1.0bfbe8p-15f;
    float kf = rintf(x * log2e);
    int k = static_cast<int>(kf);
    float r = (x - kf * ln2_hi) - kf * ln2_lo;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int k1 = k >> 1;
    int k2 = k - k1;
    return (p * pow2i(k1)) * pow2i(k2);
}

GSN_HD float det_logf(float x) {
#pragma clang fp contract(off)
    if (x != x || x < 0.0f) {
        return bits_to_float(0x7fc00000u);
    }
    if (x == 0.0f) {
        return bits_to_float(0xff800000u);
    }
    if (x == bits_to_float(0x7f800000u)) {
        return x;
    }
    int e = 0;
    if (x < 0x1p-126f) {
        x = x * 0x1p23f;
        e = -23;
    }
    uint32_t b = float_to_bits(x);
    e += static_cast<int>((b >> 23) & 0xffu) - 127;
    float m = bits_to_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421353816986083984375f) {
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float q = 1.0f / 9.0f;
    q = q * z + 1.0f / 7.0f;
    q = q * z + 1.0f / 5.0f;
    q = q * z + 1.0f / 3.0f;
    float two_s = 2.0f * s;
    float log_m = two_s + two_s * (z * q);
    float ef = static_cast<float>(e);
    const float ln2_hi = 0x1.62ep-1f;
    const float ln2_lo = 0x1.0bfbe8p-15f;
    return ef * ln2_hi + (ef * ln2_lo + log_m);
}

GSN_HD void det_sincosf(float x, float* sin_out, float* cos_out) {
#pragma clang fp contract(off)
    if (!(x - x == 0.0f)) {
        *sin_out = bits_to_float(0x7fc00000u);
        *cos_out = bits_to_float(0x7fc00000u);
        return;
    }
    const float two_over_pi = 0x1.45f306p-1f;
    const float pio2_1 = 0x1.92p+0f;
    const float pio2_2 = 0x1.fb4p-12f;
    const float pio2_3 = 0x1.4442d2p-24f;
    float kf = rintf(x * two_over_pi);
    float r = ((x - kf * pio2_1) - kf * pio2_2) - kf * pio2_3;
    int q = static_cast<int>(kf) & 3;
    float z = r * r;
    float ps = -1.0f / 5040.0f + z * (1.0f / 362880.0f);
    ps = 1.0f / 120.0f + z * ps;
    ps = -1.0f / 6.0f + z * ps;
    float s = r + (r * z) * ps;
    float pc = 1.0f / 40320.0f + z * (-1.0f / 3628800.0f);
    pc = -1.0f / 720.0f + z * pc;
    pc = 1.0f / 24.0f + z * pc;
    pc = -0.5f + z * pc;
    float c = 1.0f + z * pc;
    float so, co;
    if (q == 0) {
        so = s;
        co = c;
    }
    else if (q == 1) {
        so = c;
        co = -s;
    }
    else if (q == 2) {
        so = -s;
        co = -c;
    }
    else {
        so = -c;
        co = s;
    }
    *sin_out = so;
    *cos_out = co;
}

GSN_HD float det_sinf(float x) {
    float s, c;
    det_sincosf(x, &s, &c);
    return s;
}

// Axis-angle -> rotation matrix, row-major (reference renderer.cu:35-66).
GSN_HD void rotation_from_axis_angle(float rx, float ry, float rz, float R[9]) {
#pragma clang fp contract(off)
    float angle = sqrtf(rx * rx + ry * ry + rz * rz);
    if (angle < 1e-8f) {
        R[0] = 1.0f; R[1] = 0.0f; R[2] = 0.0f;
        R[3] = 0.0f; R[4] = 1.0f; R[5] = 0.0f;
        R[6] = 0.0f; R[7] = 0.0f; R[8] = 1.0f;
        return;
    }
    float x = rx / angle;
    float y = ry / angle;
    float z = rz / angle;
    float s, c;
    det_sincosf(angle, &s, &c);
    float t = 1.0f - c;
    R[0] = c + t * x * x;
    R[1] = t * x * y - s * z;
    R[2] = t * x * z + s * y;
    R[3] = t * x * y + s * z;
    R[4] = c + t * y * y;
    R[5] = t * y * z - s * x;
    R[6] = t * x * z - s * y;
    R[7] = t * y * z + s * x;
    R[8] = c + t * z * z;
}

GSN_HD int clamp_sh_degree(int d) {
    return d < 0 ? 0 : (d > 3 ? 3 : d);
}

// XCD-aware remap of a linear workgroup index. Workgroups are dispatched round-robin over the
// 8 XCDs (workgroup b runs on XCD b % 8); this returns an index such that each XCD receives a
// contiguous range of work items, so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    constexpr int kXcds = 8;
    int per = n / kXcds;
    int extra = n % kXcds;
    int xcd = b % kXcds;
    int idx = b / kXcds;
    // XCDs [0, extra) own per+1 items, the rest own per.
    int base = xcd < extra ? xcd * (per + 1) : extra * (per + 1) + (xcd - extra) * per;
    return base + idx;
}

} // namespace gsn
