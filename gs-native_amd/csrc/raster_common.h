// Shared pieces of the tile rasteriser forward and backward kernels.
//
// Work decomposition (gfx950, wave64): one 64-lane workgroup per 16x16 tile. Lane l owns pixel
// column l % 16 and, in each of four 4-row bands b = 0..3, the pixel at row 4*b + l / 16. A
// batch of 64 records is loaded one per lane; the per-splat values are then broadcast to the
// wave with v_readlane (scalar operands, no LDS round trip and no workgroup barrier). For each
// splat a 4-bit band mask marks the bands its alpha >= 1/255 ellipse can reach, so whole bands
// are skipped with a uniform branch. The mask is conservative (margins below), so every skipped
// (pixel, splat) pair is one the per-pixel test would have rejected anyway: results are identical
// to evaluating every pair.
#pragma once

#include "gsn_internal.h"

namespace gsn {

constexpr float kAlphaMin = 1.0f / 255.0f;
constexpr float kAlphaMax = 0.99f;
constexpr float kTransmittanceMin = 1e-4f;

__device__ __forceinline__ float readlane_f(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

__device__ __forceinline__ int readlane_i(int v, int lane) {
    return __builtin_amdgcn_readlane(v, lane);
}

// Bands (bit b = rows [y0 + 4b, y0 + 4b + 3]) of a 16x16 tile at (x0, y0) that the splat's
// alpha >= 1/255 region can touch. The region is {d : q(d) = d^T Q d <= 2 tau}, tau = ln(255 o),
// whose half extents are sqrt(2 tau * Sigma_xx) and sqrt(2 tau * Sigma_yy) with Sigma = Q^-1.
// Margins: tau + 0.25 (28% on the alpha threshold); 2 tau is further raised by 1e-5 times the
// magnitude of the three terms of q at the extents, which bounds the FP32 cancellation error of
// the per-pixel q for elongated splats (~10 ulp of the largest term); extents get 2% + 0.5 px.
// Per band, the region's x-range is taken over the band's rows only (the slab's extremes of the
// ellipse's chords; the ellipse's own extreme point when it lies inside the slab), with the same
// 2% + 0.5 px margin and never wider than the whole-tile test: a band is kept only if that range
// meets the tile's 16 columns. Non-finite inputs keep every band.
__device__ __forceinline__ int band_mask(float mx, float my, float4 conic, int x0, int y0) {
    float a = conic.x, b = conic.y, c = conic.z, o = conic.w;
    if (!(o >= kAlphaMin)) {
        return (o == o) ? 0 : 0xf;
    }
    float det = a * c - b * b;
    if (!(det > 0.0f) || !isfinite(det) || !isfinite(mx) || !isfinite(my)) {
        return 0xf;
    }
    float tau = logf(255.0f * o) + 0.25f;
    float two_tau = 2.0f * tau;
    float sxx = c / det;
    float syy = a / det;
    float ex2 = two_tau * sxx;
    float ey2 = two_tau * syy;
    float term_bound = (fabsf(a) + fabsf(b)) * ex2 + (fabsf(b) + fabsf(c)) * ey2;
    two_tau += 2e-5f * term_bound;
    const float ex_raw = sqrtf(two_tau * sxx);
    const float ey_raw = sqrtf(two_tau * syy);
    float ext_x = ex_raw * 1.02f + 0.5f;
    float ext_y = ey_raw * 1.02f + 0.5f;
    if (!isfinite(ext_x) || !isfinite(ext_y)) {
        return 0xf;
    }
    const float left = static_cast<float>(x0);
    const float right = static_cast<float>(x0 + 15);
    if (mx + ext_x < left || mx - ext_x > right) {
        return 0;
    }
    // Offsets d = (splat - pixel); the region's chord at row offset dy spans
    // dx in [(-b dy - w) / a, (-b dy + w) / a], w = sqrt(2 tau a - det dy^2). Its largest dx is
    // ex_raw at dy = -b ex_raw / c, its smallest -ex_raw at dy = b ex_raw / c; the chord ends
    // are concave / convex in dy, so over a slab without that point the extreme is at an end.
    const float inverse_a = 1.0f / a;
    const float dy_of_max = -b * ex_raw / c;
    const float two_tau_a = two_tau * a;
    const float x_margin = 0.02f * ex_raw + 0.5f;
    int mask = 0;
#pragma unroll
    for (int band = 0; band < 4; ++band) {
        float lo = static_cast<float>(y0 + 4 * band);
        float hi = lo + 3.0f;
        if (!(my + ext_y >= lo && my - ext_y <= hi)) {
            continue;
        }
        // Row offsets of the band's pixels, clamped to the region's raw y-extent.
        const float s0 = fmaxf(my - hi - 0.05f, -ey_raw);
        const float s1 = fminf(my - lo + 0.05f, ey_raw);
        bool keep = true;
        if (s0 <= s1) {
            const float w0 = sqrtf(fmaxf(two_tau_a - det * s0 * s0, 0.0f));
            const float w1 = sqrtf(fmaxf(two_tau_a - det * s1 * s1, 0.0f));
            float dx_max = (dy_of_max >= s0 && dy_of_max <= s1) ? ex_raw : fmaxf(-b * s0 + w0, -b * s1 + w1) * inverse_a;
            float dx_min = (-dy_of_max >= s0 && -dy_of_max <= s1) ? -ex_raw
                                                                  : fminf(-b * s0 - w0, -b * s1 - w1) * inverse_a;
            dx_max = fminf(dx_max + x_margin, ext_x);
            dx_min = fmaxf(dx_min - x_margin, -ext_x);
            // Pixel columns mx - dx for dx in [dx_min, dx_max] against the tile's columns.
            if (isfinite(dx_max) && isfinite(dx_min)) {
                keep = mx - dx_max <= right && mx - dx_min >= left;
            }
        }
        if (keep) {
            mask |= 1 << band;
        }
    }
    return mask;
}

} // namespace gsn

namespace gsn {

template<int Ctrl>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __builtin_amdgcn_update_dpp(0.0f, v, Ctrl, 0xf, 0xf, false);
}

// Wave-wide sums of twelve per-lane values with a transposing butterfly:
//   1. v_permlane32_swap pairs field k with k+6 across the two 32-lane halves (6 swaps),
//   2. v_permlane16_swap pairs k with k+3 across 16-lane rows (3 swaps),
//   3. four DPP butterflies (xor 1, xor 2, half-row mirror, row mirror) finish each row.
// Afterwards every lane of row r = lane / 16 holds out[k] = sum over all 64 lanes of field
// 3r + k (k = 0..2). This costs 9 swaps + 12 DPP adds instead of 12 x 6 reduction steps.
__device__ __forceinline__ void wave_sum12(const float (&v)[12], float (&out)[3]) {
    float half[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        auto swapped = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[k]), __float_as_uint(v[k + 6]), false, false);
        half[k] = __uint_as_float(swapped[0]) + __uint_as_float(swapped[1]);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        auto swapped = __builtin_amdgcn_permlane16_swap(__float_as_uint(half[k]), __float_as_uint(half[k + 3]), false, false);
        float s = __uint_as_float(swapped[0]) + __uint_as_float(swapped[1]);
        s = dpp_add<0xb1>(s);   // quad_perm [1,0,3,2]
        s = dpp_add<0x4e>(s);   // quad_perm [2,3,0,1]
        s = dpp_add<0x141>(s);  // row_half_mirror
        s = dpp_add<0x140>(s);  // row_mirror
        out[k] = s;
    }
}

} // namespace gsn
