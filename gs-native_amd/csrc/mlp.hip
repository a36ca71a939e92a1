// Learned densification proposal: a 16 -> 64 -> 64 -> 3 MLP (ReLU hidden layers, logits for
// keep / clone / split) with BF16 weights and FP32 accumulation on gfx950 MFMA
// (v_mfma_f32_32x32x16_bf16). No counterpart in the reference.
//
// One 64-lane workgroup processes tiles of 32 samples. Every GEMM of a tile is a few 32x32x16
// MFMAs with both operands staged in LDS "K-contiguous" (A[m][k], B[n][k]), so each lane loads
// its 8-element operand fragment with one 16-byte LDS read:
//   A fragment: row (l & 31) of the M block, K elements 8 * (l >> 5) .. + 7 of the K chunk,
//   B fragment: row (l & 31) of the N block, same K elements,
//   D (f32x16): lane l holds column (l & 31), register v holds row 8 (v >> 2) + 4 (l >> 5) + (v & 3).
// Forward: Z1 = X W1^T + b1, A1 = relu(Z1); Z2 = A1 W2^T + b2, A2 = relu(Z2); Z3 = A2 W3^T + b3.
// Backward (softmax cross-entropy): dZ3 = w (softmax(Z3) - onehot) / sum(w); dW3 = dZ3^T A2;
// dA2 = dZ3 W3; dZ2 = dA2 [Z2 > 0]; dW2 = dZ2^T A1; dA1 = dZ2 W2; dZ1 = dA1 [Z1 > 0];
// dW1 = dZ1^T X; db = column sums of dZ. MFMA operands are BF16 (X, weights, A1, A2, dZ*),
// accumulation and the softmax are FP32. Weight gradients are accumulated per workgroup and
// reduced in a fixed order by the Adam kernel (deterministic).
#include "gsn_internal.h"

#include <cmath>

namespace gsn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 16;
constexpr int kHidden = 64;
constexpr int kClasses = 3;
// FP32 parameter block: W1[64][16], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kIn;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kClasses * kHidden;
constexpr int kParams = kB3 + kClasses;  // 5443
// BF16 weight block: W1[64][16], W2[64][64], W3 padded to [32][64].
constexpr int kBfW1 = 0;
constexpr int kBfW2 = kBfW1 + kHidden * kIn;
constexpr int kBfW3 = kBfW2 + kHidden * kHidden;
constexpr int kBfWeights = kBfW3 + 32 * kHidden;  // 7168

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int d_row(int lane, int v) {
    return ((v >> 2) << 3) + ((lane >> 5) << 2) + (v & 3);
}

__device__ __forceinline__ bf16x8 fragment(const __bf16* base, int row, int row_stride, int k0) {
    return *reinterpret_cast<const bf16x8*>(base + row * row_stride + k0);
}

struct alignas(16) Weights {
    __bf16 w1[kHidden * kIn];        // [o][k]
    __bf16 w2[kHidden * kHidden];    // [o][i]
    __bf16 w2t[kHidden * kHidden];   // [i][o]
    __bf16 w3[32 * kHidden];         // [c][i], rows 3..31 zero
    __bf16 w3t[kHidden * 16];        // [i][c], c 3..15 zero
};

struct alignas(16) Activations {
    __bf16 x[32 * kIn];         // [m][k]
    __bf16 xt[32 * 32];         // [k][m], k 16..31 zero
    __bf16 a1[32 * kHidden];    // [m][i]
    __bf16 a1t[kHidden * 32];   // [i][m]
    __bf16 a2[32 * kHidden];
    __bf16 a2t[kHidden * 32];
    __bf16 dz3[32 * 16];        // [m][c]
    __bf16 dz3t[32 * 32];       // [c][m], c 3..31 zero
    __bf16 dz2[32 * kHidden];   // [m][o]
    __bf16 dz2t[kHidden * 32];  // [o][m]
    __bf16 dz1t[kHidden * 32];  // [o][m]
    float logits[32 * 4];
};

__device__ void load_weights(Weights& s, const __bf16* __restrict__ weights) {
    const int t = threadIdx.x;
    for (int i = t; i < kHidden * kIn; i += 64) {
        s.w1[i] = weights[kBfW1 + i];
    }
    for (int i = t; i < kHidden * kHidden; i += 64) {
        __bf16 w = weights[kBfW2 + i];
        s.w2[i] = w;
        int o = i / kHidden, j = i % kHidden;
        s.w2t[j * kHidden + o] = w;
    }
    for (int i = t; i < 32 * kHidden; i += 64) {
        __bf16 w = weights[kBfW3 + i];
        s.w3[i] = w;
    }
    for (int i = t; i < kHidden * 16; i += 64) {
        int j = i / 16, c = i % 16;
        s.w3t[i] = c < kClasses ? weights[kBfW3 + c * kHidden + j] : static_cast<__bf16>(0.0f);
    }
    __syncthreads();
}

// Stages one tile of features (rows past `count` are zero) in both layouts.
__device__ void load_tile(Activations& s, const float* __restrict__ features, int64_t first, int count) {
    const int t = threadIdx.x;
    for (int i = t; i < 32 * kIn; i += 64) {
        int m = i / kIn, k = i % kIn;
        float v = first + m < count ? features[(first + m) * kIn + k] : 0.0f;
        __bf16 b = static_cast<__bf16>(v);
        s.x[i] = b;
        s.xt[k * 32 + m] = b;
    }
    for (int i = t; i < 16 * 32; i += 64) {
        s.xt[16 * 32 + i] = static_cast<__bf16>(0.0f);
    }
    __syncthreads();
}

// Forward of one tile; leaves Z1/Z2 in registers (D layout) for the ReLU masks and the logits in
// LDS.
__device__ void forward_tile(Activations& s, const Weights& w, const float* __restrict__ biases, f32x16 z1[2],
    f32x16 z2[2]) {
    const int lane = threadIdx.x;
    const int n = lane & 31;
    const int k8 = (lane >> 5) * 8;
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 acc = {};
        acc = mfma(fragment(s.x, n, kIn, k8), fragment(w.w1, nb * 32 + n, kIn, k8), acc);
        float bias = biases[kB1 + nb * 32 + n];
        for (int v = 0; v < 16; ++v) {
            acc[v] += bias;
            int m = d_row(lane, v);
            __bf16 a = static_cast<__bf16>(fmaxf(acc[v], 0.0f));
            s.a1[m * kHidden + nb * 32 + n] = a;
            s.a1t[(nb * 32 + n) * 32 + m] = a;
        }
        z1[nb] = acc;
    }
    __syncthreads();
    for (int nb = 0; nb < 2; ++nb) {
        f32x16 acc = {};
        for (int kc = 0; kc < 4; ++kc) {
            acc = mfma(fragment(s.a1, n, kHidden, kc * 16 + k8), fragment(w.w2, nb * 32 + n, kHidden, kc * 16 + k8), acc);
        }
        float bias = biases[kB2 + nb * 32 + n];
        for (int v = 0; v < 16; ++v) {
            acc[v] += bias;
            int m = d_row(lane, v);
            __bf16 a = static_cast<__bf16>(fmaxf(acc[v], 0.0f));
            s.a2[m * kHidden + nb * 32 + n] = a;
            s.a2t[(nb * 32 + n) * 32 + m] = a;
        }
        z2[nb] = acc;
    }
    __syncthreads();
    f32x16 acc = {};
    for (int kc = 0; kc < 4; ++kc) {
        acc = mfma(fragment(s.a2, n, kHidden, kc * 16 + k8), fragment(w.w3, n, kHidden, kc * 16 + k8), acc);
    }
    if (n < kClasses) {
        float bias = biases[kB3 + n];
        for (int v = 0; v < 16; ++v) {
            s.logits[d_row(lane, v) * 4 + n] = acc[v] + bias;
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(64) void mlp_forward_kernel(const float* __restrict__ features, int count,
    const __bf16* __restrict__ weights, const float* __restrict__ biases, float* __restrict__ logits) {
    __shared__ Weights w;
    __shared__ Activations s;
    load_weights(w, weights);
    const int tiles = (count + 31) / 32;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        int64_t first = static_cast<int64_t>(tile) * 32;
        load_tile(s, features, first, count);
        f32x16 z1[2], z2[2];
        forward_tile(s, w, biases, z1, z2);
        for (int i = threadIdx.x; i < 32 * kClasses; i += 64) {
            int m = i / kClasses, c = i % kClasses;
            if (first + m < count) {
                logits[(first + m) * kClasses + c] = s.logits[m * 4 + c];
            }
        }
        __syncthreads();
    }
}

// Forward + backward over the workgroup's tiles; writes this workgroup's gradient sums
// (parameter layout) and loss sum to partials[blockIdx.x].
__global__ __launch_bounds__(64) void mlp_gradient_kernel(const float* __restrict__ features,
    const int* __restrict__ labels, const float* __restrict__ sample_weights, int count, float weight_scale,
    const __bf16* __restrict__ weights, const float* __restrict__ biases, float* __restrict__ partials,
    float* __restrict__ loss_partials) {
    __shared__ Weights w;
    __shared__ Activations s;
    load_weights(w, weights);
    const int lane = threadIdx.x;
    const int n = lane & 31;
    const int k8 = (lane >> 5) * 8;
    f32x16 gw1[2] = {}, gw2[4] = {}, gw3[2] = {};
    float gb1[2] = {0.0f, 0.0f}, gb2[2] = {0.0f, 0.0f}, gb3 = 0.0f;
    float loss = 0.0f;
    const int tiles = (count + 31) / 32;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        int64_t first = static_cast<int64_t>(tile) * 32;
        load_tile(s, features, first, count);
        f32x16 z1[2], z2[2];
        forward_tile(s, w, biases, z1, z2);
        // Softmax cross-entropy, one lane per sample.
        if (lane < 32) {
            int m = lane;
            bool valid = first + m < count;
            float weight = valid ? (sample_weights ? sample_weights[first + m] : 1.0f) * weight_scale : 0.0f;
            int label = valid ? labels[first + m] : 0;
            float l0 = s.logits[m * 4], l1 = s.logits[m * 4 + 1], l2 = s.logits[m * 4 + 2];
            float top = fmaxf(l0, fmaxf(l1, l2));
            float e0 = __expf(l0 - top), e1 = __expf(l1 - top), e2 = __expf(l2 - top);
            float sum = e0 + e1 + e2;
            float p[3] = {e0 / sum, e1 / sum, e2 / sum};
            float chosen = label == 0 ? l0 : (label == 1 ? l1 : l2);
            if (valid) {
                loss += weight * (top + __logf(sum) - chosen);
            }
            for (int c = 0; c < 16; ++c) {
                float g = c < kClasses ? weight * (p[c] - (c == label ? 1.0f : 0.0f)) : 0.0f;
                __bf16 b = static_cast<__bf16>(g);
                s.dz3[m * 16 + c] = b;
            }
            for (int c = 0; c < 32; ++c) {
                s.dz3t[c * 32 + m] = c < kClasses ? s.dz3[m * 16 + c] : static_cast<__bf16>(0.0f);
            }
        }
        __syncthreads();
        // db3, dW3 (M = class, N = hidden, K = samples).
        if (lane < kClasses) {
            for (int m = 0; m < 32; ++m) {
                gb3 += static_cast<float>(s.dz3[m * 16 + lane]);
            }
        }
        for (int nb = 0; nb < 2; ++nb) {
            for (int kc = 0; kc < 2; ++kc) {
                gw3[nb] = mfma(fragment(s.dz3t, n, 32, kc * 16 + k8), fragment(s.a2t, nb * 32 + n, 32, kc * 16 + k8),
                    gw3[nb]);
            }
        }
        // dA2 = dZ3 W3 (M = samples, N = hidden, K = class padded to 16), masked by Z2 > 0.
        for (int nb = 0; nb < 2; ++nb) {
            f32x16 acc = {};
            acc = mfma(fragment(s.dz3, n, 16, k8), fragment(w.w3t, nb * 32 + n, 16, k8), acc);
            float column = 0.0f;
            for (int v = 0; v < 16; ++v) {
                float g = z2[nb][v] > 0.0f ? acc[v] : 0.0f;
                __bf16 b = static_cast<__bf16>(g);
                int m = d_row(lane, v);
                s.dz2[m * kHidden + nb * 32 + n] = b;
                s.dz2t[(nb * 32 + n) * 32 + m] = b;
                column += static_cast<float>(b);
            }
            gb2[nb] += column;
        }
        __syncthreads();
        // dW2 (M = out, N = in, K = samples).
        for (int mb = 0; mb < 2; ++mb) {
            for (int nb = 0; nb < 2; ++nb) {
                for (int kc = 0; kc < 2; ++kc) {
                    gw2[mb * 2 + nb] = mfma(fragment(s.dz2t, mb * 32 + n, 32, kc * 16 + k8),
                        fragment(s.a1t, nb * 32 + n, 32, kc * 16 + k8), gw2[mb * 2 + nb]);
                }
            }
        }
        // dA1 = dZ2 W2 (M = samples, N = in, K = out), masked by Z1 > 0.
        for (int nb = 0; nb < 2; ++nb) {
            f32x16 acc = {};
            for (int kc = 0; kc < 4; ++kc) {
                acc = mfma(fragment(s.dz2, n, kHidden, kc * 16 + k8), fragment(w.w2t, nb * 32 + n, kHidden, kc * 16 + k8),
                    acc);
            }
            float column = 0.0f;
            for (int v = 0; v < 16; ++v) {
                float g = z1[nb][v] > 0.0f ? acc[v] : 0.0f;
                __bf16 b = static_cast<__bf16>(g);
                s.dz1t[(nb * 32 + n) * 32 + d_row(lane, v)] = b;
                column += static_cast<float>(b);
            }
            gb1[nb] += column;
        }
        __syncthreads();
        // dW1 (M = out, N = input padded to 32, K = samples).
        for (int mb = 0; mb < 2; ++mb) {
            for (int kc = 0; kc < 2; ++kc) {
                gw1[mb] = mfma(fragment(s.dz1t, mb * 32 + n, 32, kc * 16 + k8), fragment(s.xt, n, 32, kc * 16 + k8),
                    gw1[mb]);
            }
        }
        __syncthreads();
    }
    // Write this workgroup's sums. D layout: lane column n, register rows d_row(lane, v).
    float* out = partials + static_cast<int64_t>(blockIdx.x) * kParams;
    for (int mb = 0; mb < 2; ++mb) {
        for (int v = 0; v < 16; ++v) {
            int o = mb * 32 + d_row(lane, v);
            if (n < kIn) {
                out[kW1 + o * kIn + n] = gw1[mb][v];
            }
        }
    }
    for (int mb = 0; mb < 2; ++mb) {
        for (int nb = 0; nb < 2; ++nb) {
            for (int v = 0; v < 16; ++v) {
                int o = mb * 32 + d_row(lane, v);
                out[kW2 + o * kHidden + nb * 32 + n] = gw2[mb * 2 + nb][v];
            }
        }
    }
    for (int nb = 0; nb < 2; ++nb) {
        for (int v = 0; v < 16; ++v) {
            int c = d_row(lane, v);
            if (c < kClasses) {
                out[kW3 + c * kHidden + nb * 32 + n] = gw3[nb][v];
            }
        }
    }
    // Bias sums: lanes n and n + 32 hold the two row groups of the same column.
    __shared__ float bias_sums[2][2][64];
    for (int nb = 0; nb < 2; ++nb) {
        bias_sums[0][nb][lane] = gb1[nb];
        bias_sums[1][nb][lane] = gb2[nb];
    }
    __syncthreads();
    if (lane < 32) {
        for (int nb = 0; nb < 2; ++nb) {
            out[kB1 + nb * 32 + lane] = bias_sums[0][nb][lane] + bias_sums[0][nb][lane + 32];
            out[kB2 + nb * 32 + lane] = bias_sums[1][nb][lane] + bias_sums[1][nb][lane + 32];
        }
    }
    if (lane < kClasses) {
        out[kB3 + lane] = gb3;
    }
    __shared__ float losses[64];
    losses[lane] = loss;
    __syncthreads();
    if (lane == 0) {
        float total = 0.0f;
        for (int i = 0; i < 32; ++i) {
            total += losses[i];
        }
        loss_partials[blockIdx.x] = total;
    }
}

// Sums the per-workgroup gradients in a fixed order, applies Adam to the FP32 master parameters
// and refreshes the BF16 weight copy. gradients_out (optional) receives the summed gradients.
__global__ __launch_bounds__(256) void mlp_adam_kernel(const float* __restrict__ partials, int groups,
    float* __restrict__ params, float* __restrict__ first, float* __restrict__ second, __bf16* __restrict__ weights,
    float learning_rate, float first_correction, float second_correction, float* __restrict__ gradients_out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kParams) {
        return;
    }
    float g = 0.0f;
    for (int b = 0; b < groups; ++b) {
        g += partials[static_cast<int64_t>(b) * kParams + p];
    }
    if (gradients_out) {
        gradients_out[p] = g;
    }
    if (learning_rate > 0.0f) {
        float m = 0.9f * first[p] + 0.1f * g;
        float v = 0.999f * second[p] + 0.001f * g * g;
        first[p] = m;
        second[p] = v;
        params[p] -= learning_rate * (m * first_correction) / (sqrtf(v * second_correction) + 1e-8f);
    }
    float value = params[p];
    if (p < kB1) {
        weights[kBfW1 + p] = static_cast<__bf16>(value);
    }
    else if (p >= kW2 && p < kB2) {
        weights[kBfW2 + (p - kW2)] = static_cast<__bf16>(value);
    }
    else if (p >= kW3 && p < kB3) {
        weights[kBfW3 + (p - kW3)] = static_cast<__bf16>(value);
    }
}

__global__ void mlp_zero_padding_kernel(__bf16* weights) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (32 - kClasses) * kHidden) {
        weights[kBfW3 + kClasses * kHidden + i] = static_cast<__bf16>(0.0f);
    }
}

int mlp_groups(int count) {
    int tiles = (count + 31) / 32;
    return tiles < 1024 ? (tiles > 0 ? tiles : 1) : 1024;
}

} // namespace
} // namespace gsn

extern "C" int gsn_mlp_sizes(int* parameters, int* bf16_weights) {
    if (!parameters || !bf16_weights) {
        return gsn::fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_mlp_sizes: null output");
    }
    *parameters = gsn::kParams;
    *bf16_weights = gsn::kBfWeights;
    return GSN_OK;
}

extern "C" int gsn_mlp_sync_weights(gsn_context* context, const float* params, uint16_t* weights) {
    using namespace gsn;
    if (!context || !params || !weights) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_mlp_sync_weights: null argument");
    }
    mlp_zero_padding_kernel<<<grid_for((32 - kClasses) * kHidden, 256), 256, 0, context->stream>>>(
        reinterpret_cast<__bf16*>(weights));
    GSN_TRY(check_launch("mlp_zero_padding_kernel"));
    mlp_adam_kernel<<<grid_for(kParams, 256), 256, 0, context->stream>>>(nullptr, 0, const_cast<float*>(params),
        nullptr, nullptr, reinterpret_cast<__bf16*>(weights), 0.0f, 0.0f, 0.0f, nullptr);
    return check_launch("mlp_adam_kernel");
}

extern "C" int gsn_mlp_forward(gsn_context* context, const float* features, int count, const uint16_t* weights,
    const float* params, float* logits) {
    using namespace gsn;
    if (!context || count < 0 || (count > 0 && (!features || !weights || !params || !logits))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_mlp_forward: invalid arguments");
    }
    if (count == 0) {
        return GSN_OK;
    }
    mlp_forward_kernel<<<mlp_groups(count), 64, 0, context->stream>>>(features, count,
        reinterpret_cast<const __bf16*>(weights), params, logits);
    return check_launch("mlp_forward_kernel");
}

extern "C" int gsn_mlp_train_step(gsn_context* context, const float* features, const int32_t* labels,
    const float* sample_weights, int count, float* params, float* first, float* second, uint16_t* weights,
    float learning_rate, int step, float* gradients_out, double* loss_out) {
    using namespace gsn;
    if (!context || count <= 0 || !features || !labels || !params || !weights || step < 1 ||
        !(learning_rate >= 0.0f) || (learning_rate > 0.0f && (!first || !second))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_mlp_train_step: invalid arguments");
    }
    int groups = mlp_groups(count);
    float* partials = static_cast<float*>(scratch(context, kSlotMisc, static_cast<size_t>(groups) * (kParams + 1) * 4));
    if (!partials) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    float* loss_partials = partials + static_cast<size_t>(groups) * kParams;
    float weight_scale = 1.0f / static_cast<float>(count);
    mlp_gradient_kernel<<<groups, 64, 0, context->stream>>>(features, labels, sample_weights, count, weight_scale,
        reinterpret_cast<const __bf16*>(weights), params, partials, loss_partials);
    GSN_TRY(check_launch("mlp_gradient_kernel"));
    float c1 = static_cast<float>(1.0 / (1.0 - std::pow(0.9, step)));
    float c2 = static_cast<float>(1.0 / (1.0 - std::pow(0.999, step)));
    mlp_adam_kernel<<<grid_for(kParams, 256), 256, 0, context->stream>>>(partials, groups, params, first, second,
        reinterpret_cast<__bf16*>(weights), learning_rate, c1, c2, gradients_out);
    GSN_TRY(check_launch("mlp_adam_kernel"));
    if (loss_out) {
        float* host = new float[groups];
        GSN_HIP(hipMemcpyAsync(host, loss_partials, groups * sizeof(float), hipMemcpyDeviceToHost, context->stream));
        GSN_HIP(hipStreamSynchronize(context->stream));
        double total = 0.0;
        for (int b = 0; b < groups; ++b) {
            total += host[b];
        }
        delete[] host;
        *loss_out = total;
    }
    return GSN_OK;
}
