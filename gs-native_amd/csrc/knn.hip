// Three nearest neighbours of every point through a linear BVH (reference src/core/knn/knn.cu).
//
//   1. bounding box (block min/max, then one block),
//   2. 30-bit Morton code of each point over the box (10 bits per axis), stable radix sort,
//   3. points gathered into Morton order; leaves of 4 consecutive points; an implicit complete
//      binary tree of axis-aligned boxes built bottom-up level by level,
//   4. one thread per point (Morton order, so neighbouring lanes walk similar paths) runs a
//      depth-first search that visits the nearer child first and prunes a box only when its
//      lower-bound distance is strictly greater than the current third-best distance.
// Candidates are ordered by (squared distance, original index), so the three neighbours are
// unique even with duplicate points and equal the exhaustive search with the same tie rule.
// Squared distances are (dx*dx + dy*dy) + dz*dz with FP contraction off, as in the host
// restatement. Box lower bounds use the same operation order on per-axis gaps, and FP rounding
// is monotonic, so a box bound never exceeds the distance to any point inside it.
#include "gsn_internal.h"

namespace gsn {
namespace {

constexpr int kLeafSize = 4;
constexpr int kStackDepth = 64;

struct Box {
    float4 lower;  // xyz, w unused
    float4 upper;
};

__device__ __forceinline__ float ordered_min(float a, float b) {
    return a < b ? a : b;
}

__device__ __forceinline__ float ordered_max(float a, float b) {
    return a > b ? a : b;
}

__global__ __launch_bounds__(256) void knn_bounds_kernel(const float* points, int n, float* block_bounds) {
    __shared__ float s[6][256];
    int t = threadIdx.x;
    float lo[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    float hi[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int i = blockIdx.x * blockDim.x + t; i < n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float v = points[3ll * i + a];
            lo[a] = ordered_min(lo[a], v);
            hi[a] = ordered_max(hi[a], v);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a][t] = lo[a];
        s[3 + a][t] = hi[a];
    }
    __syncthreads();
    for (int stride = 128; stride > 0; stride >>= 1) {
        if (t < stride) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[a][t] = ordered_min(s[a][t], s[a][t + stride]);
                s[3 + a][t] = ordered_max(s[3 + a][t], s[3 + a][t + stride]);
            }
        }
        __syncthreads();
    }
    if (t < 6) {
        block_bounds[blockIdx.x * 6 + t] = s[t][0];
    }
}

__global__ __launch_bounds__(256) void knn_morton_kernel(const float* points, int n, const float* block_bounds,
    int bound_blocks, uint32_t* keys, uint32_t* ids) {
    __shared__ float bounds[6];
    if (threadIdx.x < 6) {
        float v = block_bounds[threadIdx.x];
        for (int b = 1; b < bound_blocks; ++b) {
            float w = block_bounds[b * 6 + threadIdx.x];
            v = threadIdx.x < 3 ? ordered_min(v, w) : ordered_max(v, w);
        }
        bounds[threadIdx.x] = v;
    }
    __syncthreads();
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) {
        return;
    }
    uint32_t key = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float extent = bounds[3 + a] - bounds[a];
        float scaled = extent > 0.0f ? (points[3ll * i + a] - bounds[a]) / extent * 1023.0f : 0.0f;
        uint32_t q = static_cast<uint32_t>(fminf(fmaxf(scaled, 0.0f), 1023.0f));
        for (int bit = 0; bit < 10; ++bit) {
            key |= ((q >> bit) & 1u) << (3 * bit + a);
        }
    }
    keys[i] = key;
    ids[i] = static_cast<uint32_t>(i);
}

__global__ __launch_bounds__(256) void knn_gather_kernel(const float* points, const uint32_t* order, int n,
    float4* sorted) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        int i = static_cast<int>(order[k]);
        sorted[k] = make_float4(points[3ll * i], points[3ll * i + 1], points[3ll * i + 2], __int_as_float(i));
    }
}

__global__ __launch_bounds__(256) void knn_leaves_kernel(const float4* sorted, int n, int leaves, Box* nodes) {
    int leaf = blockIdx.x * blockDim.x + threadIdx.x;
    if (leaf >= leaves) {
        return;
    }
    Box box;
    box.lower = make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.0f);
    box.upper = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
    for (int k = leaf * kLeafSize; k < (leaf + 1) * kLeafSize && k < n; ++k) {
        float4 p = sorted[k];
        box.lower = make_float4(ordered_min(box.lower.x, p.x), ordered_min(box.lower.y, p.y),
            ordered_min(box.lower.z, p.z), 0.0f);
        box.upper = make_float4(ordered_max(box.upper.x, p.x), ordered_max(box.upper.y, p.y),
            ordered_max(box.upper.z, p.z), 0.0f);
    }
    nodes[leaves + leaf] = box;
}

__global__ __launch_bounds__(256) void knn_level_kernel(Box* nodes, int begin) {
    int node = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= 2 * begin) {
        return;
    }
    Box a = nodes[2 * node], b = nodes[2 * node + 1];
    Box box;
    box.lower = make_float4(ordered_min(a.lower.x, b.lower.x), ordered_min(a.lower.y, b.lower.y),
        ordered_min(a.lower.z, b.lower.z), 0.0f);
    box.upper = make_float4(ordered_max(a.upper.x, b.upper.x), ordered_max(a.upper.y, b.upper.y),
        ordered_max(a.upper.z, b.upper.z), 0.0f);
    nodes[node] = box;
}

__device__ __forceinline__ float box_distance(float4 p, const Box& box) {
#pragma clang fp contract(off)
    float gx = fmaxf(fmaxf(box.lower.x - p.x, p.x - box.upper.x), 0.0f);
    float gy = fmaxf(fmaxf(box.lower.y - p.y, p.y - box.upper.y), 0.0f);
    float gz = fmaxf(fmaxf(box.lower.z - p.z, p.z - box.upper.z), 0.0f);
    return gx * gx + gy * gy + gz * gz;
}

__device__ __forceinline__ float point_distance(float4 p, float4 q) {
#pragma clang fp contract(off)
    float dx = p.x - q.x;
    float dy = p.y - q.y;
    float dz = p.z - q.z;
    return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ bool before(float d, int i, float bd, int bi) {
    return d < bd || (d == bd && i < bi);
}

__global__ __launch_bounds__(128) void knn_query_kernel(const float4* sorted, const Box* nodes, int n, int leaves,
    int* neighbors, float* mean_dist2) {
#pragma clang fp contract(off)
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) {
        return;
    }
    const float4 p = sorted[k];
    const int self = __float_as_int(p.w);
    const float inf = __int_as_float(0x7f800000);
    float bd[3] = {inf, inf, inf};
    int bi[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int stack[kStackDepth];
    int size = 0;
    int node = 1;
    while (true) {
        if (node >= leaves) {
            int leaf = node - leaves;
            for (int q = leaf * kLeafSize; q < (leaf + 1) * kLeafSize && q < n; ++q) {
                if (q == k) {
                    continue;
                }
                float4 other = sorted[q];
                float d = point_distance(p, other);
                int id = __float_as_int(other.w);
                if (before(d, id, bd[2], bi[2])) {
                    if (before(d, id, bd[1], bi[1])) {
                        bd[2] = bd[1];
                        bi[2] = bi[1];
                        if (before(d, id, bd[0], bi[0])) {
                            bd[1] = bd[0];
                            bi[1] = bi[0];
                            bd[0] = d;
                            bi[0] = id;
                        }
                        else {
                            bd[1] = d;
                            bi[1] = id;
                        }
                    }
                    else {
                        bd[2] = d;
                        bi[2] = id;
                    }
                }
            }
        }
        else {
            int left = 2 * node, right = left + 1;
            float dl = box_distance(p, nodes[left]);
            float dr = box_distance(p, nodes[right]);
            int near = dl <= dr ? left : right;
            int far = dl <= dr ? right : left;
            float dnear = dl <= dr ? dl : dr;
            float dfar = dl <= dr ? dr : dl;
            if (dfar <= bd[2] && size < kStackDepth) {
                stack[size++] = far;
            }
            if (dnear <= bd[2]) {
                node = near;
                continue;
            }
        }
        // Pop the next box that can still hold a candidate.
        node = 0;
        while (size > 0) {
            int candidate = stack[--size];
            if (box_distance(p, nodes[candidate]) <= bd[2]) {
                node = candidate;
                break;
            }
        }
        if (node == 0) {
            break;
        }
    }
    int others = n - 1;
    float result;
    if (others >= 3) {
        result = (bd[0] + bd[1] + bd[2]) / 3.0f;
    }
    else {
        float sum = bd[0];
        if (others == 2) {
            sum += bd[1];
        }
        result = sum / static_cast<float>(others);
    }
    mean_dist2[self] = result;
    if (neighbors) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            neighbors[3ll * self + j] = j < others ? bi[j] : -1;
        }
    }
}

__global__ void knn_small_kernel(int n, int* neighbors, float* mean_dist2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        mean_dist2[i] = 0.0f;
        if (neighbors) {
            neighbors[3ll * i] = -1;
            neighbors[3ll * i + 1] = -1;
            neighbors[3ll * i + 2] = -1;
        }
    }
}

} // namespace
} // namespace gsn

extern "C" int gsn_knn3(gsn_context* context, const float* points, int n, int32_t* neighbors, float* mean_dist2) {
    using namespace gsn;
    if (!context || n < 0 || (n > 0 && (!points || !mean_dist2))) {
        return fail(GSN_ERROR_INVALID_ARGUMENT, "gsn_knn3: invalid arguments");
    }
    if (n == 0) {
        return GSN_OK;
    }
    hipStream_t stream = context->stream;
    if (n < 2) {
        knn_small_kernel<<<1, 64, 0, stream>>>(n, neighbors, mean_dist2);
        return check_launch("knn_small_kernel");
    }
    int leaves = 1;
    while (leaves * kLeafSize < n) {
        leaves *= 2;
    }
    int bound_blocks = static_cast<int>(grid_for(n, 256));
    if (bound_blocks > 256) {
        bound_blocks = 256;
    }
    size_t count = static_cast<size_t>(n);
    float* bounds = static_cast<float*>(scratch(context, kSlotReduce, static_cast<size_t>(bound_blocks) * 6 * 4));
    uint32_t* keys = static_cast<uint32_t*>(scratch(context, kSlotKnnKeys, count * 4));
    uint32_t* ids = static_cast<uint32_t*>(scratch(context, kSlotKnnIds, count * 4));
    uint32_t* order = static_cast<uint32_t*>(scratch(context, kSlotKnnOrder, count * 4));
    float4* sorted = static_cast<float4*>(scratch(context, kSlotKnnPoints, count * sizeof(float4)));
    Box* nodes = static_cast<Box*>(scratch(context, kSlotKnnNodes, static_cast<size_t>(2 * leaves) * sizeof(Box)));
    if (!bounds || !keys || !ids || !order || !sorted || !nodes) {
        return GSN_ERROR_OUT_OF_MEMORY;
    }
    StageScope scope(context, kStageKnn);
    knn_bounds_kernel<<<bound_blocks, 256, 0, stream>>>(points, n, bounds);
    GSN_TRY(check_launch("knn_bounds_kernel"));
    knn_morton_kernel<<<grid_for(n, 256), 256, 0, stream>>>(points, n, bounds, bound_blocks, keys, ids);
    GSN_TRY(check_launch("knn_morton_kernel"));
    GSN_TRY(radix_sort_pairs(context, keys, ids, n, 30, nullptr, order));
    knn_gather_kernel<<<grid_for(n, 256), 256, 0, stream>>>(points, order, n, sorted);
    GSN_TRY(check_launch("knn_gather_kernel"));
    knn_leaves_kernel<<<grid_for(leaves, 256), 256, 0, stream>>>(sorted, n, leaves, nodes);
    GSN_TRY(check_launch("knn_leaves_kernel"));
    for (int begin = leaves / 2; begin > 0; begin /= 2) {
        knn_level_kernel<<<grid_for(begin, 256), 256, 0, stream>>>(nodes, begin);
        GSN_TRY(check_launch("knn_level_kernel"));
    }
    knn_query_kernel<<<grid_for(n, 128), 128, 0, stream>>>(sorted, nodes, n, leaves, neighbors, mean_dist2);
    return check_launch("knn_query_kernel");
}
