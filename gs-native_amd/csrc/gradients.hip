// Per-Gaussian gradient pass: reduces the backward partials of each Gaussian (contiguous, fixed
// order), forms the projection statistics, applies the screen-to-parameter chain rule and either
// stores the 58 parameter gradients or feeds them straight into Adam.
//
// Follows reference src/core/solver/solver.cu:
//   CollectProjectionStatisticsKernel 468-490, MapGaussianGradientsKernel<D> 492-623
//   (ComputeEquationRotation 185-221, ComputeEquationSh<D> 223-323),
//   UpdateGaussiansKernel 651-689, RemapGaussianHistoryKernel 625-649.
#include "projection.h"
#include "raster_common.h"

#include <hip/hip_ext.h>

#include <cmath>

namespace gsn {
namespace {

struct Mapped {
    float grad[10];
    float x, y, z;  // unit view direction (SH basis argument)
    bool color_ok[3];
};

// Real SH basis function k (degree <= 3) at the unit direction (x, y, z); constants as the
// reference's ComputeEquationSh (solver.cu:223-323). k is a compile-time constant after unrolling.
// No FP contraction: otherwise the fused multiply-adds chosen for the polynomials depend on the
// calling kernel, and the fused passes must give bit-identical SH gradients.
__device__ __forceinline__ float sh_basis_value(int k, float x, float y, float z) {
#pragma clang fp contract(off)
    switch (k) {
    case 0: return kShC0;
    case 1: return -kShC1 * y;
    case 2: return kShC1 * z;
    case 3: return -kShC1 * x;
    case 4: return kShC2_0 * x * y;
    case 5: return -kShC2_0 * y * z;
    case 6: return kShC2_2 * (2.0f * z * z - x * x - y * y);
    case 7: return -kShC2_0 * x * z;
    case 8: return kShC2_4 * (x * x - y * y);
    case 9: return kShC3_0 * y * (3.0f * x * x - y * y);
    case 10: return kShC3_1 * x * y * z;
    case 11: return kShC3_2 * y * (4.0f * z * z - x * x - y * y);
    case 12: return kShC3_3 * z * (2.0f * z * z - 3.0f * x * x - 3.0f * y * y);
    case 13: return kShC3_2 * x * (4.0f * z * z - x * x - y * y);
    case 14: return kShC3_5 * z * (x * x - y * y);
    default: return kShC3_0 * x * (x * x - 3.0f * y * y);
    }
}

// Partial derivatives of sh_basis_value(k, .) with respect to x, y and z (before projecting out
// the radial part).
__device__ __forceinline__ void sh_basis_derivative(int k, float x, float y, float z, float (&d)[3]) {
    constexpr float c = kShC2_0, dd = kShC2_2, e = kShC2_4;
    constexpr float a3 = kShC3_0, b3 = kShC3_1, c3 = kShC3_2, d3 = kShC3_3, e3 = kShC3_5;
    d[0] = d[1] = d[2] = 0.0f;
    switch (k) {
    case 1: d[1] = -kShC1; break;
    case 2: d[2] = kShC1; break;
    case 3: d[0] = -kShC1; break;
    case 4: d[0] = c * y; d[1] = c * x; break;
    case 5: d[1] = -c * z; d[2] = -c * y; break;
    case 6: d[0] = -2.0f * dd * x; d[1] = -2.0f * dd * y; d[2] = 4.0f * dd * z; break;
    case 7: d[0] = -c * z; d[2] = -c * x; break;
    case 8: d[0] = 2.0f * e * x; d[1] = -2.0f * e * y; break;
    case 9: d[0] = 6.0f * a3 * x * y; d[1] = 3.0f * a3 * (x * x - y * y); break;
    case 10: d[0] = b3 * y * z; d[1] = b3 * x * z; d[2] = b3 * x * y; break;
    case 11: d[0] = -2.0f * c3 * x * y; d[1] = c3 * (4.0f * z * z - x * x - 3.0f * y * y); d[2] = 8.0f * c3 * y * z; break;
    case 12: d[0] = -6.0f * d3 * x * z; d[1] = -6.0f * d3 * y * z; d[2] = d3 * (6.0f * z * z - 3.0f * x * x - 3.0f * y * y); break;
    case 13: d[0] = c3 * (4.0f * z * z - 3.0f * x * x - y * y); d[1] = -2.0f * c3 * x * y; d[2] = 8.0f * c3 * x * z; break;
    case 14: d[0] = 2.0f * e3 * x * z; d[1] = -2.0f * e3 * y * z; d[2] = e3 * (x * x - y * y); break;
    case 15: d[0] = 3.0f * a3 * (x * x - y * y); d[1] = -6.0f * a3 * x * y; break;
    default: break;
    }
}

// Chain rule from the 9 screen-space gradients to position, log-scale, axis-angle and opacity
// gradients; SH gradients are in[6 + c] * basis[k] for colour channels with value >= 0.
template<int Degree>
__device__ __forceinline__ void map_gaussian(const float* __restrict__ params, const Camera& cam, int id, int64_t n,
    const float (&in)[11], Mapped& out) {
    constexpr int s = (Degree + 1) * (Degree + 1);
    float omega[3] = {params[(kRotationIndex + 0) * n + id], params[(kRotationIndex + 1) * n + id],
        params[(kRotationIndex + 2) * n + id]};
    float rotation[9];
    rotation_from_axis_angle(omega[0], omega[1], omega[2], rotation);
    float position[3], camera_rotation[9], scale_squared[3], world[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        world[r] = params[(kPositionIndex + r) * n + id];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        position[r] = cam.t_c_w[r];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            position[r] += cam.R_c_w[r * 3 + c] * world[c];
            camera_rotation[r * 3 + c] = cam.R_c_w[r * 3] * rotation[c] + cam.R_c_w[r * 3 + 1] * rotation[3 + c] +
                cam.R_c_w[r * 3 + 2] * rotation[6 + c];
        }
        float scale = __expf(params[(kLogScaleIndex + r) * n + id]);
        scale_squared[r] = scale * scale;
    }
    float covariance[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float sum = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                sum += camera_rotation[r * 3 + j] * scale_squared[j] * camera_rotation[c * 3 + j];
            }
            covariance[r * 3 + c] = sum;
        }
    }
    float inverse_z = 1.0f / position[2];
    float nx = position[0] * inverse_z, ny = position[1] * inverse_z;
    float limit_x = kFovClampScale * 0.5f * static_cast<float>(cam.width) / cam.fx;
    float limit_y = kFovClampScale * 0.5f * static_cast<float>(cam.height) / cam.fy;
    float jx = fminf(fmaxf(nx, -limit_x), limit_x);
    float jy = fminf(fmaxf(ny, -limit_y), limit_y);
    float projection[6] = {cam.fx * inverse_z, 0.0f, -cam.fx * jx * inverse_z, 0.0f, cam.fy * inverse_z,
        -cam.fy * jy * inverse_z};
    float product[6], local_projection[6];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float p = 0.0f, l = 0.0f;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                p += projection[r * 3 + j] * covariance[j * 3 + c];
                l += projection[r * 3 + j] * camera_rotation[j * 3 + c];
            }
            product[r * 3 + c] = p;
            local_projection[r * 3 + c] = l;
        }
    }

    // SH colour (unclamped) and its derivative with respect to the world position, one basis
    // function at a time so that no basis or derivative table stays live.
    float direction[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        direction[a] = world[a];
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            direction[a] += cam.R_c_w[b * 3 + a] * cam.t_c_w[b];
        }
    }
    float inverse_length =
        1.0f / sqrtf(direction[0] * direction[0] + direction[1] * direction[1] + direction[2] * direction[2]);
    float x = direction[0] * inverse_length;
    float y = direction[1] * inverse_length;
    float z = direction[2] * inverse_length;
    float value[3] = {0.0f, 0.0f, 0.0f};
    float derivative[3][3] = {};
#pragma unroll
    for (int j = 0; j < s; ++j) {
        float basis = sh_basis_value(j, x, y, z);
        float d[3];
        sh_basis_derivative(j, x, y, z, d);
#pragma unroll
        for (int channel = 0; channel < 3; ++channel) {
            float coefficient = params[(kShIndex + channel * 16 + j) * n + id];
            value[channel] += coefficient * basis;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                derivative[channel][a] += coefficient * d[a];
            }
        }
    }
    out.x = x;
    out.y = y;
    out.z = z;
    float color_derivatives[9];
#pragma unroll
    for (int channel = 0; channel < 3; ++channel) {
        out.color_ok[channel] = value[channel] + 0.5f >= 0.0f;
        float radial = derivative[channel][0] * x + derivative[channel][1] * y + derivative[channel][2] * z;
        color_derivatives[channel * 3 + 0] = (derivative[channel][0] - radial * x) * inverse_length;
        color_derivatives[channel * 3 + 1] = (derivative[channel][1] - radial * y) * inverse_length;
        color_derivatives[channel * 3 + 2] = (derivative[channel][2] - radial * z) * inverse_length;
    }

    float right_gradient[3];
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        float dx = cam.R_c_w[axis], dy = cam.R_c_w[3 + axis], dz = cam.R_c_w[6 + axis];
        float du = (dx - nx * dz) * inverse_z;
        float dv = (dy - ny * dz) * inverse_z;
        float djx = nx >= -limit_x && nx <= limit_x ? du : 0.0f;
        float djy = ny >= -limit_y && ny <= limit_y ? dv : 0.0f;
        float delta[6] = {-cam.fx * dz * inverse_z * inverse_z, 0.0f, -cam.fx * (djx - jx * dz * inverse_z) * inverse_z,
            0.0f, -cam.fy * dz * inverse_z * inverse_z, -cam.fy * (djy - jy * dz * inverse_z) * inverse_z};
        float daa = 0.0f, dab = 0.0f, ddd = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            daa += 2.0f * delta[j] * product[j];
            dab += delta[j] * product[3 + j] + delta[3 + j] * product[j];
            ddd += 2.0f * delta[3 + j] * product[3 + j];
        }
        float center_gradient =
            in[0] * cam.fx * du + in[1] * cam.fy * dv + in[2] * daa + in[3] * dab + in[4] * ddd;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if (out.color_ok[c]) {
                center_gradient += in[6 + c] * color_derivatives[c * 3 + axis];
            }
        }
        out.grad[kPositionIndex + axis] = center_gradient;
        float a = local_projection[axis], b = local_projection[3 + axis];
        out.grad[kLogScaleIndex + axis] = 2.0f * scale_squared[axis] * (in[2] * a * a + in[3] * a * b + in[4] * b * b);
        int j = (axis + 1) % 3, k = (axis + 2) % 3;
        float difference = scale_squared[j] - scale_squared[k];
        right_gradient[axis] = difference *
            (2.0f * in[2] * local_projection[j] * local_projection[k] +
                in[3] * (local_projection[j] * local_projection[3 + k] + local_projection[k] * local_projection[3 + j]) +
                2.0f * in[4] * local_projection[3 + j] * local_projection[3 + k]);
    }
    float theta_squared = omega[0] * omega[0] + omega[1] * omega[1] + omega[2] * omega[2];
    float sa = 0.5f, sb = 1.0f / 6.0f;
    if (theta_squared > 1e-4f) {
        float theta = sqrtf(theta_squared);
        float half_sinc = det_sinf(0.5f * theta) / (0.5f * theta);
        sa = 0.5f * half_sinc * half_sinc;
        sb = (theta - det_sinf(theta)) / (theta_squared * theta);
    }
    else {
        sa -= theta_squared / 24.0f;
        sb -= theta_squared / 120.0f;
    }
    float cross[3];
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        int j = (axis + 1) % 3, k = (axis + 2) % 3;
        cross[axis] = omega[j] * right_gradient[k] - omega[k] * right_gradient[j];
    }
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        int j = (axis + 1) % 3, k = (axis + 2) % 3;
        out.grad[kRotationIndex + axis] =
            right_gradient[axis] + sa * cross[axis] + sb * (omega[j] * cross[k] - omega[k] * cross[j]);
    }
    out.grad[kOpacityIndex] = in[5];
}

// Dense Adam step of one entry (reference solver.cu:666-688); no FP contraction so that the
// result equals the host restatement bit for bit for equal inputs.
__device__ __forceinline__ float learning_rate_for_row(int row, float position_learning_rate) {
    if (row < kLogScaleIndex) {
        return position_learning_rate;
    }
    if (row < kRotationIndex) {
        return 0.005f;
    }
    if (row < kOpacityIndex) {
        return 0.001f;
    }
    if (row < kShIndex) {
        return 0.025f;
    }
    return (row - kShIndex) % 16 == 0 ? 0.0025f : 0.0025f / 20.0f;
}

__device__ __forceinline__ void adam_values(float& param, float& first, float& second, float gradient,
    float learning_rate, float first_correction, float second_correction) {
#pragma clang fp contract(off)
    float m = 0.9f * first + 0.1f * gradient;
    float v = 0.999f * second + 0.001f * gradient * gradient;
    first = m;
    second = v;
    param -= learning_rate * (m * first_correction) / (sqrtf(v * second_correction) + 1e-15f);
}

// Two-float vector for non-temporal moment pairs (the builtins take native vector types).
typedef float moment_pair __attribute__((ext_vector_type(2)));

__device__ __forceinline__ moment_pair load_moments(const float* moments, int64_t index) {
    return __builtin_nontemporal_load(reinterpret_cast<const moment_pair*>(moments + index));
}

__device__ __forceinline__ void store_moments(float* moments, int64_t index, float m, float v) {
    moment_pair pair = {m, v};
    __builtin_nontemporal_store(pair, reinterpret_cast<moment_pair*>(moments + index));
}

constexpr int kAdamGroup = 16;

constexpr int kGradientBlock = kGradientGroup;
constexpr int kChunks = 2;  // 64-record chunks per wave pass
// A Gaussian owning at least this many stamped records of a 64-record chunk has them summed by the
// whole wave (wave_sum12) instead of one by one by its own lane.
constexpr int kWideOwner = 16;

// Density-control accumulators that a pass may add its statistics to (all null when it does not):
// the same additions as gsn_density_accumulate makes from the written statistics.
struct StatisticSums {
    float* sums;
    float* absolute_sums;
    uint32_t* counts;
};

// Projection statistics of one Gaussian (reference CollectProjectionStatisticsKernel,
// solver.cu:468-490): pixel-space norms of the mean gradient and of its absolute counterpart, and
// whether the Gaussian had records. No FP contraction, so that every kernel writes identical bits.
__device__ __forceinline__ void write_statistics(const Camera& cam, const float (&in)[11], int id, int64_t n,
    int begin, int end, float* statistics, float* absolute_statistics, const StatisticSums& accumulate) {
#pragma clang fp contract(off)
    float x = in[0] * (0.5f * static_cast<float>(cam.width));
    float y = in[1] * (0.5f * static_cast<float>(cam.height));
    float ax = in[9] * (0.5f * static_cast<float>(cam.width));
    float ay = in[10] * (0.5f * static_cast<float>(cam.height));
    const float norm = sqrtf(x * x + y * y);
    const float visible = end > begin ? 1.0f : 0.0f;
    const float absolute = sqrtf(ax * ax + ay * ay);
    statistics[id] = norm;
    statistics[n + id] = visible;
    absolute_statistics[id] = absolute;
    absolute_statistics[n + id] = in[10];
    if (accumulate.sums) {
        accumulate.sums[id] += norm;
        accumulate.absolute_sums[id] += absolute;
        accumulate.counts[id] += static_cast<uint32_t>(visible);
    }
}

// Sums of the stamped partials of Gaussian `id` (Gaussians at or past `limit` own no records);
// all 64 lanes of the wave call it together. slots: the wave's LDS staging area,
// kChunks * 64 * 3 float4.
// Each Gaussian owns the contiguous run of partial records of its tile rectangle, so a wave's 64
// Gaussians own one contiguous run. The backward wrote a partial only for the records its tile
// reached before every pixel was done, and stamped those records' flag bytes with this step's
// epoch (at 1080p with a converged scene about 4% of the records). The wave walks its run
// kChunks x 64 records at a time, one record per lane and chunk: the flag bytes are read
// together, the stamped partials are loaded (coalesced, neighbouring lanes hold neighbouring
// records) into the wave's LDS slots, and each lane then adds the stamped records of its own
// Gaussian from LDS in record order; a Gaussian owning kWideOwner or more of a chunk's stamped
// records instead adds the chunk's wave-wide sum of them. Every order is fixed, so results are
// identical run to run. Work per wave
// follows the length of its run, not the largest rectangle in it, and each pass costs two
// dependent global accesses (flags, then partials) for kChunks x 64 records.
__device__ __forceinline__ void reduce_records(const float* __restrict__ partials,
    const int* __restrict__ record_offsets, const uint8_t* __restrict__ record_flags, uint8_t epoch, int id, int limit,
    int record_count, float4* slots, float (&in)[11], int& begin, int& end) {
    const int lane = threadIdx.x & 63;
    begin = record_count;
    end = record_count;
    if (id < limit) {
        begin = min(record_offsets[id], record_count);
        end = min(record_offsets[id + 1], record_count);
    }
    const int last_lane = min(63, max(0, limit - 1 - (id - lane)));
    const int wave_begin = __shfl(begin, 0);
    const int wave_end = __shfl(end, last_lane);
#pragma unroll
    for (int k = 0; k < 11; ++k) {
        in[k] = 0.0f;
    }
    const float4* records = reinterpret_cast<const float4*>(partials);
    // Flags of the first pass; each pass then loads the next pass's flags while its partials are
    // in flight, so a pass waits on memory about once. Flag reads past the run clamp to its last
    // record (they are never used).
    static_assert(kChunks == 2, "the staging below is written out for two chunks");
    uint8_t flag0 = 0, flag1 = 0;
    if (wave_begin < wave_end) {
        flag0 = record_flags[min(wave_begin + lane, wave_end - 1)];
        flag1 = record_flags[min(wave_begin + 64 + lane, wave_end - 1)];
    }
    for (int pass = wave_begin; pass < wave_end; pass += kChunks * 64) {
        bool reached[kChunks];
        reached[0] = pass + lane < wave_end && flag0 == epoch;
        reached[1] = pass + 64 + lane < wave_end && flag1 == epoch;
        uint64_t ballots[kChunks];
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            ballots[c] = __ballot(reached[c]);
        }
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float4 a0 = zero, b0 = zero, c0 = zero, a1 = zero, b1 = zero, c1 = zero;
        const int64_t source = 3 * static_cast<int64_t>(pass + lane);
        if (reached[0]) {
            a0 = records[source];
            b0 = records[source + 1];
            c0 = records[source + 2];
        }
        if (reached[1]) {
            a1 = records[source + 3 * 64];
            b1 = records[source + 3 * 64 + 1];
            c1 = records[source + 3 * 64 + 2];
        }
        const int next = pass + kChunks * 64;
        if (next < wave_end) {
            flag0 = record_flags[min(next + lane, wave_end - 1)];
            flag1 = record_flags[min(next + 64 + lane, wave_end - 1)];
        }
        if (reached[0]) {
            slots[3 * lane] = a0;
            slots[3 * lane + 1] = b0;
            slots[3 * lane + 2] = c0;
        }
        if (reached[1]) {
            slots[3 * (64 + lane)] = a1;
            slots[3 * (64 + lane) + 1] = b1;
            slots[3 * (64 + lane) + 2] = c1;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int c = 0; c < kChunks; ++c) {
            // This lane's records [begin, end) within the chunk.
            const int chunk = pass + c * 64;
            const int low = min(max(begin - chunk, 0), 64);
            const int high = min(max(end - chunk, 0), 64);
            uint64_t mine = ballots[c];
            mine &= high >= 64 ? ~0ull : ((1ull << high) - 1);
            mine &= low >= 64 ? 0ull : ~((1ull << low) - 1);
            // Wide owners: each lane takes its own record position if the owner stamped it there,
            // the wave sums the twelve fields and the owner adds the chunk's total. A single new
            // Gaussian can own thousands of records (a screen-filling split), which one lane would
            // otherwise add one LDS round trip at a time while its wave holds the pass open.
            const uint64_t wide = __ballot(__popcll(mine) >= kWideOwner);
            if (wide) {
                const int own = 3 * (c * 64 + lane);
                const float4 a = slots[own], b = slots[own + 1], cc = slots[own + 2];
                uint64_t rest = wide;
                while (rest) {
                    const int j = __builtin_ctzll(rest);
                    rest &= rest - 1;
                    const uint32_t owned_lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(mine), j);
                    const uint32_t owned_hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(mine >> 32), j);
                    const bool take = (((static_cast<uint64_t>(owned_hi) << 32) | owned_lo) >> lane) & 1;
                    const float v[12] = {take ? a.x : 0.0f, take ? a.y : 0.0f, take ? a.z : 0.0f,
                        take ? a.w : 0.0f, take ? b.x : 0.0f, take ? b.y : 0.0f, take ? b.z : 0.0f,
                        take ? b.w : 0.0f, take ? cc.x : 0.0f, take ? cc.y : 0.0f, take ? cc.z : 0.0f, 0.0f};
                    float out[3];
                    wave_sum12(v, out);
                    float total[12];
#pragma unroll
                    for (int row = 0; row < 4; ++row) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            total[3 * row + k] = __uint_as_float(
                                __builtin_amdgcn_readlane(__float_as_uint(out[k]), 16 * row));
                        }
                    }
                    if (lane == j) {
#pragma unroll
                        for (int k = 0; k < 11; ++k) {
                            in[k] += total[k];
                        }
                    }
                }
                if (__popcll(mine) >= kWideOwner) {
                    mine = 0;
                }
            }
            while (mine) {
                const int k = __builtin_ctzll(mine);
                mine &= mine - 1;
                const float4 a = slots[3 * (c * 64 + k)], b = slots[3 * (c * 64 + k) + 1],
                             cc = slots[3 * (c * 64 + k) + 2];
                in[0] += a.x;
                in[1] += a.y;
                in[2] += a.z;
                in[3] += a.w;
                in[4] += b.x;
                in[5] += b.y;
                in[6] += b.z;
                in[7] += b.w;
                in[8] += cc.x;
                in[9] += cc.y;
                in[10] += cc.z;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// The next view's projection outputs, written by the fused pass from the updated parameters.
struct NextProjection {
    Camera camera;
    float2* means;
    float4* conic_opacities;
    float* colors;
    float* depths;
    int4* tile_rects;
    int* record_counts;
};

// Project: with Fused, also project each Gaussian's updated parameters for the next view
// (project_geometry after the first Adam group, which holds rows 0-15; each SH coefficient's
// colour term as its row is updated, in coefficient order), so the next iteration reads no
// parameter rows to project. The outputs equal the projection kernel's bit for bit.
template<int Degree, bool Fused, bool Streaming, bool Project>
__global__ __launch_bounds__(kGradientBlock) void gaussian_gradients_kernel(const int* __restrict__ group_order,
    const float* __restrict__ partials,
    const int* __restrict__ record_offsets, const uint8_t* __restrict__ record_flags, uint8_t epoch, int count,
    int record_count, Camera cam, float* params, float* gradients, float* screen, float* statistics,
    float* absolute_statistics, float* moments, float position_learning_rate, float first_correction,
    float second_correction, NextProjection next, StatisticSums accumulate) {
    static_assert(Fused || !Project, "the next view is projected from the Adam update");
    __shared__ float4 s_records[kGradientBlock / 64][kChunks * 64 * 3];
    const int wave = threadIdx.x >> 6;
    // Workgroups run heaviest record run first (group_order); results do not depend on the order.
    const int group = group_order ? group_order[blockIdx.x] : static_cast<int>(blockIdx.x);
    const int id = group * kGradientBlock + threadIdx.x;
    const int64_t n = count;
    if (Project && id == 0) {
        next.record_counts[count] = 0;
    }
    float in[11];
    int begin, end;
    reduce_records(partials, record_offsets, record_flags, epoch, id, count, record_count, s_records[wave], in, begin,
        end);
    if (id >= count) {
        return;
    }
    write_statistics(cam, in, id, n, begin, end, statistics, absolute_statistics, accumulate);
    if (screen) {
#pragma unroll
        for (int k = 0; k < 11; ++k) {
            screen[k * n + id] = in[k];
        }
    }
    bool active = false;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        active = active || in[k] != 0.0f;
    }
    Mapped mapped;
    if (active) {
        map_gaussian<Degree>(params, cam, id, n, in, mapped);
    }
    constexpr int coefficients = (Degree + 1) * (Degree + 1);
    auto row_gradient = [&](int row) {
        float g = 0.0f;
        if (active) {
            if (row < kShIndex) {
                g = mapped.grad[row];
            }
            else {
                int channel = (row - kShIndex) / 16;
                int k = (row - kShIndex) % 16;
                g = (k < coefficients && mapped.color_ok[channel])
                    ? in[6 + channel] * sh_basis_value(k, mapped.x, mapped.y, mapped.z) : 0.0f;
            }
        }
        return g;
    };
    if (Fused) {
        // Adam over the 58 rows in groups: all loads of a group are issued before its stores, so
        // each wave keeps 2 * kAdamGroup loads in flight instead of one row's. The moments are
        // blocked (a wave's 64 pairs of a row are one aligned 512-byte piece), so they always
        // stream with non-temporal accesses that keep them from displacing cached data. The
        // parameter rows are the interface's [58][N] layout: they stream non-temporally only when
        // every row starts on a 128-byte line (Streaming); off the line grid (a count that is not
        // a multiple of 32, as after density control) each wave's row segment straddles three
        // lines, which ordinary accesses share with the neighbouring wave through the cache.
        float* __restrict__ p = params;
        float* __restrict__ mv = moments;
        constexpr int coeffs = (Degree + 1) * (Degree + 1);
        float basis[kShCoeffCount];
        float colour[3] = {0.0f, 0.0f, 0.0f};
        bool visible = false;
#pragma unroll
        for (int base = 0; base < kParamCount; base += kAdamGroup) {
            float pv[kAdamGroup];
            moment_pair qv[kAdamGroup];
#pragma unroll
            for (int k = 0; k < kAdamGroup; ++k) {
                if (base + k < kParamCount) {
                    int64_t index = (base + k) * n + id;
                    pv[k] = Streaming ? __builtin_nontemporal_load(p + index) : p[index];
                    qv[k] = load_moments(mv, moment_index(id, base + k));
                }
            }
#pragma unroll
            for (int k = 0; k < kAdamGroup; ++k) {
                if (base + k < kParamCount) {
                    int row = base + k;
                    int64_t index = row * n + id;
                    float m = qv[k].x, v = qv[k].y;
                    adam_values(pv[k], m, v, row_gradient(row), learning_rate_for_row(row, position_learning_rate),
                        first_correction, second_correction);
                    if (Streaming) {
                        __builtin_nontemporal_store(pv[k], p + index);
                    }
                    else {
                        p[index] = pv[k];
                    }
                    store_moments(mv, moment_index(id, row), m, v);
                }
            }
            if (Project) {
                static_assert(kAdamGroup >= kShIndex, "the first group holds the geometry rows");
                if (base == 0) {
                    next.record_counts[id] = 0;
                    visible = project_geometry<Degree>([&](int r) { return pv[r]; }, next.camera, id, next.means,
                        next.conic_opacities, next.depths, next.tile_rects, next.record_counts, basis);
                }
#pragma unroll
                for (int k = 0; k < kAdamGroup; ++k) {
                    const int row = base + k;
                    if (row >= kShIndex && row < kParamCount && (row - kShIndex) % kShCoeffCount < coeffs) {
                        const int channel = (row - kShIndex) / kShCoeffCount;
                        colour[channel] = add_sh_term(colour[channel], basis[(row - kShIndex) % kShCoeffCount], pv[k]);
                    }
                }
            }
        }
        if (Project && visible) {
#pragma unroll
            for (int channel = 0; channel < 3; ++channel) {
                next.colors[static_cast<int64_t>(id) * 3 + channel] = finish_colour(colour[channel]);
            }
        }
    }
    else if (gradients) {
#pragma unroll
        for (int row = 0; row < kParamCount; ++row) {
            gradients[row * n + id] = row_gradient(row);
        }
    }
}

__global__ __launch_bounds__(256) void adam_kernel(const float* __restrict__ gradients, float* params, float* moments,
    int count, float position_learning_rate, float first_correction, float second_correction) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    int row = blockIdx.y;
    int64_t index = static_cast<int64_t>(row) * count + id;
    int64_t pair = moment_index(id, row);
    float p = params[index];
    moment_pair q = load_moments(moments, pair);
    float m = q.x, v = q.y;
    adam_values(p, m, v, gradients[index], learning_rate_for_row(row, position_learning_rate), first_correction,
        second_correction);
    params[index] = p;
    store_moments(moments, pair, m, v);
}

// One thread per new Gaussian: its parent id is read once and its 58 moment pairs are gathered
// kRemapRows rows at a time (all loads of a group in flight); neighbouring new ids mostly have
// neighbouring parents, so each row's gathers and stores stay within one or two 512-byte pieces.
constexpr int kRemapRows = 8;

__global__ __launch_bounds__(256) void remap_history_kernel(const int* parent_ids, const float* previous, float* moments,
    int previous_count, int count) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    const int parent = parent_ids ? parent_ids[id] : -1;
    const bool inherited = parent >= 0 && parent < previous_count;
#pragma unroll
    for (int base = 0; base < kParamCount; base += kRemapRows) {
        float2 pairs[kRemapRows];
#pragma unroll
        for (int k = 0; k < kRemapRows; ++k) {
            const int row = base + k;
            pairs[k] = make_float2(0.0f, 0.0f);
            if (row < kParamCount && inherited) {
                pairs[k] = *reinterpret_cast<const float2*>(previous + moment_index(parent, row));
            }
        }
#pragma unroll
        for (int k = 0; k < kRemapRows; ++k) {
            const int row = base + k;
            if (row < kParamCount) {
                *reinterpret_cast<float2*>(moments + moment_index(id, row)) = pairs[k];
            }
        }
    }
}

// Blocked moments <-> two [58][count] arrays (first and second moments).
__global__ __launch_bounds__(256) void moments_unpack_kernel(const float* moments, float* first, float* second,
    int count) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    int row = blockIdx.y;
    float2 pair = *reinterpret_cast<const float2*>(moments + moment_index(id, row));
    first[static_cast<int64_t>(row) * count + id] = pair.x;
    second[static_cast<int64_t>(row) * count + id] = pair.y;
}

__global__ __launch_bounds__(256) void moments_pack_kernel(const float* first, const float* second, float* moments,
    int count) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= count) {
        return;
    }
    int row = blockIdx.y;
    *reinterpret_cast<float2*>(moments + moment_index(id, row)) =
        make_float2(first[static_cast<int64_t>(row) * count + id], second[static_cast<int64_t>(row) * count + id]);
}

// Zeroes the moment pairs of rows [row, row + rows) for all `count` Gaussians.
__global__ __launch_bounds__(256) void zero_rows_kernel(float* moments, int count, int row, int rows) {
    int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < static_cast<int64_t>(count) * rows) {
        int r = row + static_cast<int>(i / count);
        int64_t id = i % count;
        *reinterpret_cast<float2*>(moments + moment_index(id, r)) = make_float2(0.0f, 0.0f);
    }
}

// start / stop: the timing events of the stage, recorded by the dispatch itself (null when the
// stage is not timed).
template<int Degree, bool Fused, bool Project>
void launch_gradients_degree(bool streaming, unsigned grid, hipStream_t stream, hipEvent_t start, hipEvent_t stop,
    const int* group_order, const float* partials, const int* record_offsets, const uint8_t* record_flags,
    uint8_t epoch, int count, int record_count, const Camera& cam, float* params, float* gradients, float* screen,
    float* statistics, float* absolute_statistics, float* moments, float lr, float c1, float c2,
    const NextProjection& next, const StatisticSums& accumulate) {
    if (Fused && streaming) {
        hipExtLaunchKernelGGL(gaussian_gradients_kernel<Degree, Fused, Fused, Project>, dim3(grid), dim3(kGradientBlock),
            0, stream, start, stop, 0, group_order, partials, record_offsets, record_flags, epoch, count, record_count,
            cam, params, gradients, screen, statistics, absolute_statistics, moments, lr, c1, c2, next, accumulate);
    }
    else {
        hipExtLaunchKernelGGL(gaussian_gradients_kernel<Degree, Fused, false, Project>, dim3(grid), dim3(kGradientBlock),
            0, stream, start, stop, 0, group_order, partials, record_offsets, record_flags, epoch, count, record_count,
            cam, params, gradients, screen, statistics, absolute_statistics, moments, lr, c1, c2, next, accumulate);
    }
}

// next: with Fused, the next view's projection to write from the updated parameters (or null).
// accumulate: density-control accumulators to add the statistics to (or null).
template<bool Fused>
int launch_gradients(gsn_context* context, const Camera& cam, int count, int record_count, int sh_degree,
    const int* group_order, const float* partials, const int* record_offsets, const uint8_t* record_flags,
    uint8_t epoch, float* params, float* gradients, float* screen, float* statistics, float* absolute_statistics,
    float* moments, float lr, float c1, float c2, const NextProjection* next = nullptr,
    const StatisticSums* accumulate = nullptr) {
    unsigned grid = grid_for(count, kGradientBlock);
    hipStream_t stream = context->stream;
    // Non-temporal streaming of the parameter rows only when all of them start on a 128-byte line
    // (the blocked moments always stream non-temporally).
    const bool streaming = Fused && rows_line_aligned(params, count);
    const NextProjection none{};
    const NextProjection& target = next ? *next : none;
    const StatisticSums no_sums{};
    const StatisticSums& sums = accumulate ? *accumulate : no_sums;
    StageScope scope(context, kStageGaussianGradients, true);
    hipEvent_t start = scope.start_event();
    hipEvent_t stop = scope.stop_event();
#define GSN_LAUNCH_DEGREE(D)                                                                                           \
    if (Fused && next) {                                                                                               \
        launch_gradients_degree<D, Fused, Fused>(streaming, grid, stream, start, stop, group_order, partials,          \
            record_offsets, record_flags, epoch, count, record_count, cam, params, gradients, screen, statistics,       \
            absolute_statistics, moments, lr, c1, c2, target, sums);                                                   \
    }                                                                                                                  \
    else {                                                                                                             \
        launch_gradients_degree<D, Fused, false>(streaming, grid, stream, start, stop, group_order, partials,          \
            record_offsets, record_flags, epoch, count, record_count, cam, params, gradients, screen, statistics,       \
            absolute_statistics, moments, lr, c1, c2, target, sums);                                                   \
    }
    switch (clamp_sh_degree(sh_degree)) {
    case 0:
        GSN_LAUNCH_DEGREE(0)
        break;
    case 1:
        GSN_LAUNCH_DEGREE(1)
        break;
    case 2:
        GSN_LAUNCH_DEGREE(2)
        break;
    default:
        GSN_LAUNCH_DEGREE(3)
        break;
    }
#undef GSN_LAUNCH_DEGREE
    return check_launch("gaussian_gradients_kernel");
}

int validate_gradient_inputs(const gsn_context* context, const gsn_camera* camera, int count, const void* params,
    const int* ids, const int64_t* tile_offsets, int record_count, const int* record_offsets, const int* rects,
    const float* means, const float* conics, const float* colors, const float* image, const float* image_gradients,
    const float* statistics, const float* absolute_statistics) {
    if (!context || !camera || count < 0 || record_count < 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gaussian gradients: invalid arguments");
    }
    if (camera->width <= 0 || camera->height <= 0 || camera->tile_count_x <= 0 || camera->tile_count_y <= 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gaussian gradients: invalid camera");
    }
    if (count > 0 && (!params || !tile_offsets || !record_offsets || !rects || !means || !conics || !colors || !image ||
                         !image_gradients || !statistics || !absolute_statistics)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gaussian gradients: null buffer");
    }
    if (record_count > 0 && !ids) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gaussian gradients: null tile ids");
    }
    return GSN_OK;
}

} // namespace
} // namespace gsn

extern "C" int gsn_gaussian_gradients(gsn_context* context, const gsn_camera* camera, int count, int sh_degree,
    const float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets, int record_count,
    const int* record_offsets, const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* image, const float* image_gradients, float* gradients, float* screen_gradients,
    float* statistics, float* absolute_statistics) {
    using namespace gsn;
    GSN_TRY(validate_gradient_inputs(context, camera, count, params, tile_gaussian_ids, tile_offsets, record_count,
        record_offsets, tile_rects, means, conic_opacities, colors, image, image_gradients, statistics,
        absolute_statistics));
    if (count == 0) {
        return GSN_OK;
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    float* partials = static_cast<float*>(
        scratch(context, kSlotPartials, static_cast<size_t>(record_count > 0 ? record_count : 1) * kPartialFloats * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    const uint8_t* flags = nullptr;
    uint8_t epoch = 0;
    const int* group_order = nullptr;
    if (record_count > 0) {
        GSN_TRY(launch_splat_backward(context, cam, count, tile_gaussian_ids, tile_offsets, means, conic_opacities,
            colors, tile_rects, record_offsets, image, image_gradients, partials, record_count, &flags, &epoch,
            &group_order));
    }
    return launch_gradients<false>(context, cam, count, record_count, sh_degree, group_order, partials,
        record_offsets, flags, epoch, const_cast<float*>(params), gradients, screen_gradients, statistics,
        absolute_statistics, nullptr, 0.0f, 0.0f, 0.0f);
}

extern "C" int gsn_gaussian_gradients_adam(gsn_context* context, const gsn_camera* camera, int count, int sh_degree,
    float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets, int record_count,
    const int* record_offsets, const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* image, const float* image_gradients, float* moments,
    float position_learning_rate, float first_correction, float second_correction, float* statistics,
    float* absolute_statistics) {
    using namespace gsn;
    GSN_TRY(validate_gradient_inputs(context, camera, count, params, tile_gaussian_ids, tile_offsets, record_count,
        record_offsets, tile_rects, means, conic_opacities, colors, image, image_gradients, statistics,
        absolute_statistics));
    if (count == 0) {
        return GSN_OK;
    }
    if (!moments || !(position_learning_rate >= 0.0f) || !std::isfinite(position_learning_rate)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_gaussian_gradients_adam: invalid optimiser arguments");
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    float* partials = static_cast<float*>(
        scratch(context, kSlotPartials, static_cast<size_t>(record_count > 0 ? record_count : 1) * kPartialFloats * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    const uint8_t* flags = nullptr;
    uint8_t epoch = 0;
    const int* group_order = nullptr;
    if (record_count > 0) {
        GSN_TRY(launch_splat_backward(context, cam, count, tile_gaussian_ids, tile_offsets, means, conic_opacities,
            colors, tile_rects, record_offsets, image, image_gradients, partials, record_count, &flags, &epoch,
            &group_order));
    }
    return launch_gradients<true>(context, cam, count, record_count, sh_degree, group_order, partials,
        record_offsets, flags, epoch, params, nullptr, nullptr, statistics, absolute_statistics, moments,
        position_learning_rate, first_correction, second_correction);
}

extern "C" int gsn_gaussian_gradients_adam_project(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets, int record_count,
    const int* record_offsets, const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* image, const float* image_gradients, float* moments,
    float position_learning_rate, float first_correction, float second_correction, float* statistics,
    float* absolute_statistics, const gsn_camera* next_camera, float* next_means, float* next_conic_opacities,
    float* next_colors, float* next_depths, int* next_tile_rects, int* next_record_offsets, float* gradient_sums,
    float* absolute_gradient_sums, uint32_t* observation_counts) {
    using namespace gsn;
    GSN_TRY(validate_gradient_inputs(context, camera, count, params, tile_gaussian_ids, tile_offsets, record_count,
        record_offsets, tile_rects, means, conic_opacities, colors, image, image_gradients, statistics,
        absolute_statistics));
    if (!moments || !(position_learning_rate >= 0.0f) || !std::isfinite(position_learning_rate) || !next_camera ||
        !next_record_offsets || next_record_offsets == record_offsets ||
        (count > 0 && (!next_means || !next_conic_opacities || !next_colors || !next_depths || !next_tile_rects))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_gaussian_gradients_adam_project: invalid arguments");
    }
    if (next_camera->width <= 0 || next_camera->height <= 0 || next_camera->tile_count_x <= 0 ||
        next_camera->tile_count_y <= 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_gaussian_gradients_adam_project: invalid next camera");
    }
    const int accumulators = (gradient_sums != nullptr) + (absolute_gradient_sums != nullptr) +
                             (observation_counts != nullptr);
    if (accumulators != 0 && accumulators != 3) {
        return fail(GSN_ERROR_INVALID_ARGUMENT,
            "gsn_gaussian_gradients_adam_project: give all three density accumulators or none");
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    float* partials = static_cast<float*>(
        scratch(context, kSlotPartials, static_cast<size_t>(record_count > 0 ? record_count : 1) * kPartialFloats * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    const uint8_t* flags = nullptr;
    uint8_t epoch = 0;
    const int* group_order = nullptr;
    if (record_count > 0) {
        GSN_TRY(launch_splat_backward(context, cam, count, tile_gaussian_ids, tile_offsets, means, conic_opacities,
            colors, tile_rects, record_offsets, image, image_gradients, partials, record_count, &flags, &epoch,
            &group_order));
    }
    // From here on the next view's outputs are the context's last projection.
    GSN_TRY(record_projection(context, count, next_means, next_conic_opacities, next_colors, next_depths,
        next_tile_rects, next_record_offsets));
    if (count == 0) {
        return check_hip(hipMemsetAsync(next_record_offsets, 0, sizeof(int), context->stream), "hipMemsetAsync");
    }
    NextProjection next;
    memcpy(&next.camera, next_camera, sizeof(next.camera));
    next.means = reinterpret_cast<float2*>(next_means);
    next.conic_opacities = reinterpret_cast<float4*>(next_conic_opacities);
    next.colors = next_colors;
    next.depths = next_depths;
    next.tile_rects = reinterpret_cast<int4*>(next_tile_rects);
    next.record_counts = next_record_offsets;
    const StatisticSums accumulate{gradient_sums, absolute_gradient_sums, observation_counts};
    return launch_gradients<true>(context, cam, count, record_count, sh_degree, group_order, partials,
        record_offsets, flags, epoch, params, nullptr, nullptr, statistics, absolute_statistics, moments,
        position_learning_rate, first_correction, second_correction, &next, &accumulate);
}

extern "C" int gsn_gaussian_gradients_stored(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, const float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets, int record_count,
    const int* record_offsets, const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* final_transmittance, const int32_t* last_contributor, const float* background,
    const float* image_gradients, float* gradients, float* screen_gradients, float* statistics,
    float* absolute_statistics) {
    using namespace gsn;
    // The image itself is not read by this form; validation only needs a non-null stand-in.
    GSN_TRY(validate_gradient_inputs(context, camera, count, params, tile_gaussian_ids, tile_offsets, record_count,
        record_offsets, tile_rects, means, conic_opacities, colors, image_gradients, image_gradients, statistics,
        absolute_statistics));
    if (count == 0) {
        return GSN_OK;
    }
    if (!final_transmittance || !last_contributor || !background) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_gaussian_gradients_stored: null forward state");
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    float* partials = static_cast<float*>(
        scratch(context, kSlotPartials, static_cast<size_t>(record_count > 0 ? record_count : 1) * kPartialFloats * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    const uint8_t* flags = nullptr;
    uint8_t epoch = 0;
    const int* group_order = nullptr;
    if (record_count > 0) {
        StoredState stored{final_transmittance, last_contributor, {background[0], background[1], background[2]}};
        GSN_TRY(launch_splat_backward(context, cam, count, tile_gaussian_ids, tile_offsets, means, conic_opacities,
            colors, tile_rects, record_offsets, nullptr, image_gradients, partials, record_count, &flags, &epoch,
            &group_order,
            &stored));
    }
    return launch_gradients<false>(context, cam, count, record_count, sh_degree, group_order, partials,
        record_offsets, flags, epoch, const_cast<float*>(params), gradients, screen_gradients, statistics,
        absolute_statistics, nullptr, 0.0f, 0.0f, 0.0f);
}

extern "C" int gsn_gaussian_gradients_adam_stored(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets, int record_count,
    const int* record_offsets, const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* final_transmittance, const int32_t* last_contributor, const float* background,
    const float* image_gradients, float* moments, float position_learning_rate, float first_correction,
    float second_correction, float* statistics, float* absolute_statistics) {
    using namespace gsn;
    GSN_TRY(validate_gradient_inputs(context, camera, count, params, tile_gaussian_ids, tile_offsets, record_count,
        record_offsets, tile_rects, means, conic_opacities, colors, image_gradients, image_gradients, statistics,
        absolute_statistics));
    if (count == 0) {
        return GSN_OK;
    }
    if (!final_transmittance || !last_contributor || !background || !moments ||
        !(position_learning_rate >= 0.0f) || !std::isfinite(position_learning_rate)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_gaussian_gradients_adam_stored: invalid arguments");
    }
    Camera cam;
    memcpy(&cam, camera, sizeof(cam));
    float* partials = static_cast<float*>(
        scratch(context, kSlotPartials, static_cast<size_t>(record_count > 0 ? record_count : 1) * kPartialFloats * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    const uint8_t* flags = nullptr;
    uint8_t epoch = 0;
    const int* group_order = nullptr;
    if (record_count > 0) {
        StoredState stored{final_transmittance, last_contributor, {background[0], background[1], background[2]}};
        GSN_TRY(launch_splat_backward(context, cam, count, tile_gaussian_ids, tile_offsets, means, conic_opacities,
            colors, tile_rects, record_offsets, nullptr, image_gradients, partials, record_count, &flags, &epoch,
            &group_order,
            &stored));
    }
    return launch_gradients<true>(context, cam, count, record_count, sh_degree, group_order, partials,
        record_offsets, flags, epoch, params, nullptr, nullptr, statistics, absolute_statistics, moments,
        position_learning_rate, first_correction, second_correction);
}

extern "C" int gsn_adam_update(gsn_context* context, int count, const float* gradients, float* params,
    float* moments, float position_learning_rate, float first_correction, float second_correction) {
    using namespace gsn;
    if (!context || count < 0 || !(position_learning_rate >= 0.0f) || !std::isfinite(position_learning_rate)) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_adam_update: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    if (!gradients || !params || !moments) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_adam_update: null buffer");
    }
    dim3 grid(grid_for(count, 256), kParamCount);
    StageScope scope(context, kStageAdam);
    adam_kernel<<<grid, 256, 0, context->stream>>>(gradients, params, moments, count, position_learning_rate,
        first_correction, second_correction);
    return check_launch("adam_kernel");
}

extern "C" int gsn_remap_history(gsn_context* context, const int* parent_ids, int previous_count, int count,
    const float* previous_moments, float* moments) {
    using namespace gsn;
    if (!context || count < 0 || previous_count < 0) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_remap_history: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    if (!moments || (previous_count > 0 && parent_ids && !previous_moments) || previous_moments == moments) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_remap_history: null or aliased moment buffer");
    }
    StageScope scope(context, kStageDensity);
    remap_history_kernel<<<grid_for(count, 256), 256, 0, context->stream>>>(parent_ids, previous_moments, moments,
        previous_moments ? previous_count : 0, count);
    return check_launch("remap_history_kernel");
}

extern "C" int gsn_zero_rows(gsn_context* context, float* moments, int count, int row, int rows) {
    using namespace gsn;
    if (!context || count < 0 || row < 0 || rows < 0 || row + rows > kParamCount) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_zero_rows: invalid arguments");
    }
    int64_t total = static_cast<int64_t>(count) * rows;
    if (total == 0) {
        return GSN_OK;
    }
    if (!moments) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_zero_rows: null buffer");
    }
    zero_rows_kernel<<<grid_for(total, 256), 256, 0, context->stream>>>(moments, count, row, rows);
    return check_launch("zero_rows_kernel");
}

extern "C" size_t gsn_moment_floats(int count) {
    return count > 0 ? static_cast<size_t>(gsn::moment_floats(count)) : 0;
}

extern "C" int gsn_moments_unpack(gsn_context* context, int count, const float* moments, float* first,
    float* second) {
    using namespace gsn;
    if (!context || count < 0 || (count > 0 && (!moments || !first || !second))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_moments_unpack: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    dim3 grid(grid_for(count, 256), kParamCount);
    moments_unpack_kernel<<<grid, 256, 0, context->stream>>>(moments, first, second, count);
    return check_launch("moments_unpack_kernel");
}

extern "C" int gsn_moments_pack(gsn_context* context, int count, const float* first, const float* second,
    float* moments) {
    using namespace gsn;
    if (!context || count < 0 || (count > 0 && (!moments || !first || !second))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_moments_pack: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    dim3 grid(grid_for(count, 256), kParamCount);
    moments_pack_kernel<<<grid, 256, 0, context->stream>>>(first, second, moments, count);
    return check_launch("moments_pack_kernel");
}
