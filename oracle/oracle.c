/*
 * oracle.c - CPU restatement of the reference training hot path. TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker;
 * never linked into or called by the product. See oracle.h for the contract of each function.
 *
 * Build: oracle/build.sh (gcc -O2 -ffp-contract=off, no -ffast-math).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the parallel loops (OpenMP). 1 by default: the loops then run in the sequential order
 * the tests were written against. Only the CPU timing of a full iteration raises it; every loop
 * that is split keeps each output's summation order, except the splat backward, whose per-thread
 * partial sums are added in thread order (deterministic for a given thread count). */
static int g_threads = 1;

void oracle_set_threads(int threads) {
    g_threads = threads < 1 ? 1 : threads;
}

#define PARAMS 58
#define SH_INDEX 10
#define OPACITY_INDEX 9
#define ROTATION_INDEX 6
#define LOG_SCALE_INDEX 3

static const float kShC0 = 0.28209479177387814f;
static const float kShC1 = 0.4886025119029199f;
static const float kShC2[5] = {1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f,
    -1.0925484305920792f, 0.5462742152960396f};
static const float kShC3[7] = {-0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
    0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f};

/* ---- FP32 transcendentals -----------------------------------------------------------------
 * Same algorithm and operation order as the device functions (Cody-Waite reduction + Taylor
 * polynomials); with contraction off and IEEE SSE arithmetic the results are bit-identical. */

static float bits_to_float(uint32_t b) {
    float f;
    memcpy(&f, &b, 4);
    return f;
}

static uint32_t float_to_bits(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return b;
}

static float pow2i(int n) {
    return bits_to_float((uint32_t)(n + 127) << 23);
}

float oracle_expf(float x) {
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
    int k = (int)kf;
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

float oracle_logf(float x) {
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
    e += (int)((b >> 23) & 0xffu) - 127;
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
    float ef = (float)e;
    const float ln2_hi = 0x1.62ep-1f;
    const float ln2_lo = 0x1.0bfbe8p-15f;
    return ef * ln2_hi + (ef * ln2_lo + log_m);
}

void oracle_sincosf(float x, float* sin_out, float* cos_out) {
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
    int q = (int)kf & 3;
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

static float oracle_sinf(float x) {
    float s, c;
    oracle_sincosf(x, &s, &c);
    return s;
}

/* renderer.cu:35-66 */
static void rotation_f32(float rx, float ry, float rz, float R[9]) {
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
    oracle_sincosf(angle, &s, &c);
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

/* ---- projection (renderer.cu:108-221), FP32 mirror ------------------------------------- */

void oracle_project_f32(const float* params, int count, const oracle_camera* cam, int sh_degree, float* means,
    float* conic_opacities, float* colors, float* depths, int* tile_rects, int* record_counts) {
    sh_degree = sh_degree < 0 ? 0 : (sh_degree > 3 ? 3 : sh_degree);
    record_counts[count] = 0;
    const int64_t n = count;
    const float* Rc = cam->R_c_w;
    const float* tc = cam->t_c_w;
    for (int i = 0; i < count; ++i) {
        record_counts[i] = 0;
        float px0 = params[0 * n + i], py0 = params[1 * n + i], pz0 = params[2 * n + i];
        float px = Rc[0] * px0 + Rc[1] * py0 + Rc[2] * pz0 + tc[0];
        float py = Rc[3] * px0 + Rc[4] * py0 + Rc[5] * pz0 + tc[1];
        float z = Rc[6] * px0 + Rc[7] * py0 + Rc[8] * pz0 + tc[2];
        float near_depth = fmaxf(cam->near_depth, 1e-6f);
        if (z <= near_depth || z > fmaxf(cam->far_depth, near_depth)) {
            continue;
        }
        float Rg[9];
        rotation_f32(params[(ROTATION_INDEX + 0) * n + i], params[(ROTATION_INDEX + 1) * n + i],
            params[(ROTATION_INDEX + 2) * n + i], Rg);
        float scale[3] = {oracle_expf(params[(LOG_SCALE_INDEX + 0) * n + i]),
            oracle_expf(params[(LOG_SCALE_INDEX + 1) * n + i]), oracle_expf(params[(LOG_SCALE_INDEX + 2) * n + i])};
        float M[9], cov[9];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                M[r * 3 + c] = (Rc[r * 3] * Rg[c] + Rc[r * 3 + 1] * Rg[3 + c] + Rc[r * 3 + 2] * Rg[6 + c]) * scale[c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                cov[r * 3 + c] = M[r * 3] * M[c * 3] + M[r * 3 + 1] * M[c * 3 + 1] + M[r * 3 + 2] * M[c * 3 + 2];
            }
        }
        float limit_x = 1.3f * 0.5f * (float)cam->width / cam->fx;
        float limit_y = 1.3f * 0.5f * (float)cam->height / cam->fy;
        float x = fminf(fmaxf(px / z, -limit_x), limit_x) * z;
        float y = fminf(fmaxf(py / z, -limit_y), limit_y) * z;
        float J[6] = {cam->fx / z, 0.0f, -cam->fx * x / (z * z), 0.0f, cam->fy / z, -cam->fy * y / (z * z)};
        float JS[6];
        for (int r = 0; r < 2; ++r) {
            for (int c = 0; c < 3; ++c) {
                JS[r * 3 + c] = J[r * 3] * cov[c] + J[r * 3 + 1] * cov[3 + c] + J[r * 3 + 2] * cov[6 + c];
            }
        }
        float a = JS[0] * J[0] + JS[1] * J[1] + JS[2] * J[2] + 0.3f;
        float b = JS[0] * J[3] + JS[1] * J[4] + JS[2] * J[5];
        float c = JS[3] * J[3] + JS[4] * J[4] + JS[5] * J[5] + 0.3f;
        float det = a * c - b * b;
        /* Near-degenerate covariances: det and conic from the exact FP64 products. */
        int degenerate = det < 1e-3f * (a * c);
        double det64 = 0.0;
        if (degenerate) {
            det64 = (double)a * (double)c - (double)b * (double)b;
            det = (float)det64;
        }
        if (!(det > 0.0f)) {
            continue;
        }
        float middle = 0.5f * (a + c);
        float radius = ceilf(3.0f * sqrtf(middle + sqrtf(fmaxf(0.1f, middle * middle - det))));
        float u = cam->fx * px / z + cam->cx;
        float v = cam->fy * py / z + cam->cy;
        if (!isfinite(radius) || !isfinite(u) || !isfinite(v)) {
            continue;
        }
        int rx0 = (int)fminf((float)cam->tile_count_x, fmaxf(0.0f, (u - radius) / 16));
        int ry0 = (int)fminf((float)cam->tile_count_y, fmaxf(0.0f, (v - radius) / 16));
        int rx1 = (int)fminf((float)cam->tile_count_x, fmaxf(0.0f, (u + radius + 16 - 1) / 16));
        int ry1 = (int)fminf((float)cam->tile_count_y, fmaxf(0.0f, (v + radius + 16 - 1) / 16));
        if (rx0 >= rx1 || ry0 >= ry1) {
            continue;
        }
        depths[i] = z;
        tile_rects[4 * i + 0] = rx0;
        tile_rects[4 * i + 1] = ry0;
        tile_rects[4 * i + 2] = rx1;
        tile_rects[4 * i + 3] = ry1;
        record_counts[i] = (rx1 - rx0) * (ry1 - ry0);
        float cx, cy, cz;
        if (degenerate) {
            cx = (float)((double)c / det64);
            cy = (float)(-(double)b / det64);
            cz = (float)((double)a / det64);
        }
        else {
            float det_inv = 1.0f / det;
            cx = c * det_inv;
            cy = -b * det_inv;
            cz = a * det_inv;
        }
        conic_opacities[4 * i + 0] = cx;
        conic_opacities[4 * i + 1] = cy;
        conic_opacities[4 * i + 2] = cz;
        conic_opacities[4 * i + 3] = 1.0f / (1.0f + oracle_expf(-params[OPACITY_INDEX * n + i]));
        means[2 * i] = u;
        means[2 * i + 1] = v;
        float dx = px0 + Rc[0] * tc[0] + Rc[3] * tc[1] + Rc[6] * tc[2];
        float dy = py0 + Rc[1] * tc[0] + Rc[4] * tc[1] + Rc[7] * tc[2];
        float dz = pz0 + Rc[2] * tc[0] + Rc[5] * tc[1] + Rc[8] * tc[2];
        float length_inv = 1.0f / sqrtf(dx * dx + dy * dy + dz * dz);
        float xs = dx * length_inv, ys = dy * length_inv, zs = dz * length_inv;
        float basis[16];
        basis[0] = kShC0;
        if (sh_degree >= 1) {
            basis[1] = -kShC1 * ys;
            basis[2] = kShC1 * zs;
            basis[3] = -kShC1 * xs;
        }
        if (sh_degree >= 2) {
            float xx = xs * xs, yy = ys * ys, zz = zs * zs;
            float xy = xs * ys, yz = ys * zs, xz = xs * zs;
            basis[4] = kShC2[0] * xy;
            basis[5] = kShC2[1] * yz;
            basis[6] = kShC2[2] * (2.0f * zz - xx - yy);
            basis[7] = kShC2[3] * xz;
            basis[8] = kShC2[4] * (xx - yy);
            if (sh_degree >= 3) {
                basis[9] = kShC3[0] * ys * (3.0f * xx - yy);
                basis[10] = kShC3[1] * xy * zs;
                basis[11] = kShC3[2] * ys * (4.0f * zz - xx - yy);
                basis[12] = kShC3[3] * zs * (2.0f * zz - 3.0f * xx - 3.0f * yy);
                basis[13] = kShC3[4] * xs * (4.0f * zz - xx - yy);
                basis[14] = kShC3[5] * zs * (xx - yy);
                basis[15] = kShC3[6] * xs * (xx - 3.0f * yy);
            }
        }
        int coeffs = (sh_degree + 1) * (sh_degree + 1);
        for (int ch = 0; ch < 3; ++ch) {
            float value = 0.0f;
            for (int j = 0; j < coeffs; ++j) {
                value += basis[j] * params[(SH_INDEX + ch * 16 + j) * n + i];
            }
            colors[3 * i + ch] = fmaxf(value + 0.5f, 0.0f);
        }
    }
}

int oracle_scan_i32(int* values, int count) {
    int running = 0;
    for (int i = 0; i <= count; ++i) {
        int v = values[i];
        values[i] = running;
        running += v;
    }
    return values[count];
}

/* ---- tile lists (renderer.cu:223-271, 541-569) --------------------------------------------- */

typedef struct {
    uint64_t key;
    int64_t position;
    int id;
} record_t;

static int compare_records(const void* a, const void* b) {
    const record_t* x = (const record_t*)a;
    const record_t* y = (const record_t*)b;
    if (x->key != y->key) {
        return x->key < y->key ? -1 : 1;
    }
    return x->position < y->position ? -1 : (x->position > y->position ? 1 : 0);
}

void oracle_tile_lists(const oracle_camera* cam, int count, const int* tile_rects, const float* depths,
    const int* record_offsets, int record_count, int* tile_gaussian_ids, int64_t* tile_offsets, uint64_t* sorted_keys) {
    int tile_count = cam->tile_count_x * cam->tile_count_y;
    for (int t = 0; t <= tile_count; ++t) {
        tile_offsets[t] = 0;
    }
    if (record_count == 0) {
        return;
    }
    record_t* records = (record_t*)malloc(sizeof(record_t) * (size_t)record_count);
    for (int id = 0; id < count; ++id) {
        int begin = record_offsets[id];
        int total = record_offsets[id + 1] - begin;
        if (total == 0) {
            continue;
        }
        const int* rect = tile_rects + 4 * id;
        int width = rect[2] - rect[0];
        uint64_t depth = float_to_bits(depths[id]);
        for (int local = 0; local < total; ++local) {
            uint64_t tile = (uint64_t)(rect[1] + local / width) * (uint64_t)cam->tile_count_x + (uint64_t)(rect[0] + local % width);
            records[begin + local].key = (tile << 32) | depth;
            records[begin + local].position = begin + local;
            records[begin + local].id = id;
        }
    }
    qsort(records, (size_t)record_count, sizeof(record_t), compare_records);
    for (int r = 0; r < record_count; ++r) {
        tile_gaussian_ids[r] = records[r].id;
        if (sorted_keys) {
            sorted_keys[r] = records[r].key;
        }
        int tile = (int)(records[r].key >> 32);
        int previous = r == 0 ? -1 : (int)(records[r - 1].key >> 32);
        if (tile != previous) {
            for (int t = previous + 1; t <= tile; ++t) {
                tile_offsets[t] = r;
            }
        }
        if (r == record_count - 1) {
            for (int t = tile + 1; t <= tile_count; ++t) {
                tile_offsets[t] = record_count;
            }
        }
    }
    free(records);
}

/* ---- forward rasteriser (renderer.cu:273-338), FP32 mirror ------------------------------- */

static int near_threshold(double value, double threshold, double relative) {
    return fabs(value - threshold) <= relative * fabs(threshold);
}

int oracle_render_f32(const oracle_camera* cam, const int* ids, const int64_t* tile_offsets, const float* means,
    const float* conics, const float* colors, const float* background, float* image, uint8_t* ambiguous) {
    int ambiguous_count = 0;
    for (int ty = 0; ty < cam->tile_count_y; ++ty) {
        for (int tx = 0; tx < cam->tile_count_x; ++tx) {
            int tile = ty * cam->tile_count_x + tx;
            for (int ly = 0; ly < 16; ++ly) {
                for (int lx = 0; lx < 16; ++lx) {
                    int x = tx * 16 + lx, y = ty * 16 + ly;
                    if (x >= cam->width || y >= cam->height) {
                        continue;
                    }
                    float T = 1.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
                    int flag = 0;
                    for (int64_t r = tile_offsets[tile]; r < tile_offsets[tile + 1]; ++r) {
                        int id = ids[r];
                        float dx = means[2 * id] - (float)x;
                        float dy = means[2 * id + 1] - (float)y;
                        const float* q = conics + 4 * id;
                        float t1 = q[0] * dx * dx, t2 = q[2] * dy * dy, t3 = q[1] * dx * dy;
                        float power = -0.5f * (t1 + t2) - t3;
                        if (fabsf(power) <= 1e-5f * (fabsf(t1) + fabsf(t2) + 2.0f * fabsf(t3)) + 1e-30f) {
                            flag = 1;
                        }
                        if (power > 0.0f) {
                            continue;
                        }
                        float alpha = fminf(0.99f, q[3] * expf(power));
                        if (near_threshold(alpha, 1.0 / 255.0, 1e-4)) {
                            flag = 1;
                        }
                        if (alpha < 1.0f / 255.0f) {
                            continue;
                        }
                        float next = T * (1.0f - alpha);
                        if (near_threshold(next, 1e-4, 1e-4)) {
                            flag = 1;
                        }
                        if (next < 1e-4f) {
                            break;
                        }
                        float w = alpha * T;
                        cr += colors[3 * id] * w;
                        cg += colors[3 * id + 1] * w;
                        cb += colors[3 * id + 2] * w;
                        T = next;
                    }
                    float* pixel = image + ((int64_t)y * cam->width + x) * 3;
                    pixel[0] = cr + T * background[0];
                    pixel[1] = cg + T * background[1];
                    pixel[2] = cb + T * background[2];
                    if (ambiguous) {
                        ambiguous[(int64_t)y * cam->width + x] = (uint8_t)flag;
                    }
                    ambiguous_count += flag;
                }
            }
        }
    }
    return ambiguous_count;
}

/* ---- FP64 projection (renderer.cu:108-221; cpu_reference.h:47-79 with its guards) -------- */

static void rotation_f64(double rx, double ry, double rz, double R[9]) {
    double angle = sqrt(rx * rx + ry * ry + rz * rz);
    if (angle < 1e-8) {
        for (int k = 0; k < 9; ++k) {
            R[k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
        return;
    }
    double x = rx / angle, y = ry / angle, z = rz / angle;
    double c = cos(angle), s = sin(angle), t = 1.0 - c;
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

static void sh_basis_f64(double x, double y, double z, double basis[16]) {
    /* cpu_reference.h:11-21 */
    basis[0] = 0.28209479177387814;
    basis[1] = -0.4886025119029199 * y;
    basis[2] = 0.4886025119029199 * z;
    basis[3] = -0.4886025119029199 * x;
    basis[4] = 1.0925484305920792 * x * y;
    basis[5] = -1.0925484305920792 * y * z;
    basis[6] = 0.31539156525252005 * (2 * z * z - x * x - y * y);
    basis[7] = -1.0925484305920792 * x * z;
    basis[8] = 0.5462742152960396 * (x * x - y * y);
    basis[9] = -0.5900435899266435 * y * (3 * x * x - y * y);
    basis[10] = 2.890611442640554 * x * y * z;
    basis[11] = -0.4570457994644658 * y * (4 * z * z - x * x - y * y);
    basis[12] = 0.3731763325901154 * z * (2 * z * z - 3 * x * x - 3 * y * y);
    basis[13] = -0.4570457994644658 * x * (4 * z * z - x * x - y * y);
    basis[14] = 1.445305721320277 * z * (x * x - y * y);
    basis[15] = -0.5900435899266435 * x * (x * x - 3 * y * y);
}

/* Exported for the pinning tests (renderer.cu:35-66 and :68-106 as the oracle restates them). */
void oracle_rotation_f32(float rx, float ry, float rz, float* R) {
    rotation_f32(rx, ry, rz, R);
}

void oracle_rotation_f64(double rx, double ry, double rz, double* R) {
    rotation_f64(rx, ry, rz, R);
}

void oracle_sh_basis_f64(double x, double y, double z, double* basis) {
    sh_basis_f64(x, y, z, basis);
}

void oracle_project_f64(const float* params, int count, const oracle_camera* cam, int sh_degree, double* means,
    double* conics, double* colors, double* depths, int* tile_rects, uint8_t* visible) {
    sh_degree = sh_degree < 0 ? 0 : (sh_degree > 3 ? 3 : sh_degree);
    const int64_t n = count;
    double Rc[9], tc[3];
    for (int k = 0; k < 9; ++k) {
        Rc[k] = cam->R_c_w[k];
    }
    for (int k = 0; k < 3; ++k) {
        tc[k] = cam->t_c_w[k];
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int i = 0; i < count; ++i) {
        visible[i] = 0;
        double p0 = params[0 * n + i], p1 = params[1 * n + i], p2 = params[2 * n + i];
        double px = Rc[0] * p0 + Rc[1] * p1 + Rc[2] * p2 + tc[0];
        double py = Rc[3] * p0 + Rc[4] * p1 + Rc[5] * p2 + tc[1];
        double z = Rc[6] * p0 + Rc[7] * p1 + Rc[8] * p2 + tc[2];
        double near_depth = fmax(cam->near_depth, 1e-6);
        if (z <= near_depth || z > fmax(cam->far_depth, near_depth)) {
            continue;
        }
        double Rg[9];
        rotation_f64(params[(ROTATION_INDEX + 0) * n + i], params[(ROTATION_INDEX + 1) * n + i],
            params[(ROTATION_INDEX + 2) * n + i], Rg);
        double scale[3];
        for (int k = 0; k < 3; ++k) {
            scale[k] = exp((double)params[(LOG_SCALE_INDEX + k) * n + i]);
        }
        double M[9], cov[9];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                M[r * 3 + c] = (Rc[r * 3] * Rg[c] + Rc[r * 3 + 1] * Rg[3 + c] + Rc[r * 3 + 2] * Rg[6 + c]) * scale[c];
            }
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                cov[r * 3 + c] = M[r * 3] * M[c * 3] + M[r * 3 + 1] * M[c * 3 + 1] + M[r * 3 + 2] * M[c * 3 + 2];
            }
        }
        double limit_x = 0.65 * cam->width / cam->fx;
        double limit_y = 0.65 * cam->height / cam->fy;
        double x = fmin(fmax(px / z, -limit_x), limit_x) * z;
        double y = fmin(fmax(py / z, -limit_y), limit_y) * z;
        double J[6] = {cam->fx / z, 0.0, -cam->fx * x / (z * z), 0.0, cam->fy / z, -cam->fy * y / (z * z)};
        double JS[6];
        for (int r = 0; r < 2; ++r) {
            for (int c = 0; c < 3; ++c) {
                JS[r * 3 + c] = J[r * 3] * cov[c] + J[r * 3 + 1] * cov[3 + c] + J[r * 3 + 2] * cov[6 + c];
            }
        }
        double a = JS[0] * J[0] + JS[1] * J[1] + JS[2] * J[2] + 0.3;
        double b = JS[0] * J[3] + JS[1] * J[4] + JS[2] * J[5];
        double c = JS[3] * J[3] + JS[4] * J[4] + JS[5] * J[5] + 0.3;
        double det = a * c - b * b;
        if (!(det > 0.0)) {
            continue;
        }
        double middle = 0.5 * (a + c);
        double radius = ceil(3.0 * sqrt(middle + sqrt(fmax(0.1, middle * middle - det))));
        double u = cam->fx * px / z + cam->cx;
        double v = cam->fy * py / z + cam->cy;
        if (!isfinite(radius) || !isfinite(u) || !isfinite(v)) {
            continue;
        }
        int rx0 = (int)fmin(cam->tile_count_x, fmax(0.0, (u - radius) / 16));
        int ry0 = (int)fmin(cam->tile_count_y, fmax(0.0, (v - radius) / 16));
        int rx1 = (int)fmin(cam->tile_count_x, fmax(0.0, (u + radius + 15) / 16));
        int ry1 = (int)fmin(cam->tile_count_y, fmax(0.0, (v + radius + 15) / 16));
        if (rx0 >= rx1 || ry0 >= ry1) {
            continue;
        }
        visible[i] = 1;
        depths[i] = z;
        tile_rects[4 * i + 0] = rx0;
        tile_rects[4 * i + 1] = ry0;
        tile_rects[4 * i + 2] = rx1;
        tile_rects[4 * i + 3] = ry1;
        means[2 * i] = u;
        means[2 * i + 1] = v;
        conics[4 * i + 0] = c / det;
        conics[4 * i + 1] = -b / det;
        conics[4 * i + 2] = a / det;
        conics[4 * i + 3] = 1.0 / (1.0 + exp(-(double)params[OPACITY_INDEX * n + i]));
        double d0 = p0 + Rc[0] * tc[0] + Rc[3] * tc[1] + Rc[6] * tc[2];
        double d1 = p1 + Rc[1] * tc[0] + Rc[4] * tc[1] + Rc[7] * tc[2];
        double d2 = p2 + Rc[2] * tc[0] + Rc[5] * tc[1] + Rc[8] * tc[2];
        double length = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        double basis[16];
        sh_basis_f64(d0 / length, d1 / length, d2 / length, basis);
        int coeffs = (sh_degree + 1) * (sh_degree + 1);
        for (int ch = 0; ch < 3; ++ch) {
            double value = 0.5;
            for (int k = 0; k < coeffs; ++k) {
                value += basis[k] * params[(SH_INDEX + ch * 16 + k) * n + i];
            }
            colors[3 * i + ch] = fmax(value, 0.0);
        }
    }
}

/* Visible Gaussians in stable depth order (cpu_reference.h:36-41). */
typedef struct {
    double depth;
    int id;
} depth_t;

static int compare_depth(const void* a, const void* b) {
    const depth_t* x = (const depth_t*)a;
    const depth_t* y = (const depth_t*)b;
    if (x->depth != y->depth) {
        return x->depth < y->depth ? -1 : 1;
    }
    return x->id - y->id;
}

static int* depth_order(int count, const double* depths, const uint8_t* visible, int* ordered) {
    depth_t* items = (depth_t*)malloc(sizeof(depth_t) * (size_t)(count > 0 ? count : 1));
    int m = 0;
    for (int i = 0; i < count; ++i) {
        if (visible[i]) {
            items[m].depth = depths[i];
            items[m].id = i;
            ++m;
        }
    }
    qsort(items, (size_t)m, sizeof(depth_t), compare_depth);
    int* order = (int*)malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    for (int k = 0; k < m; ++k) {
        order[k] = items[k].id;
    }
    free(items);
    *ordered = m;
    return order;
}

/* Per-tile lists of depth-ordered Gaussians: offsets[T+1], ids. */
static int* tile_buckets(const oracle_camera* cam, const int* order, int ordered, const int* rects, int64_t** offsets_out) {
    int tiles = cam->tile_count_x * cam->tile_count_y;
    int64_t* offsets = (int64_t*)calloc((size_t)tiles + 1, sizeof(int64_t));
    for (int k = 0; k < ordered; ++k) {
        const int* r = rects + 4 * order[k];
        for (int ty = r[1]; ty < r[3]; ++ty) {
            for (int tx = r[0]; tx < r[2]; ++tx) {
                offsets[ty * cam->tile_count_x + tx + 1] += 1;
            }
        }
    }
    for (int t = 0; t < tiles; ++t) {
        offsets[t + 1] += offsets[t];
    }
    int64_t total = offsets[tiles];
    int* ids = (int*)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    int64_t* cursor = (int64_t*)malloc(sizeof(int64_t) * (size_t)(tiles > 0 ? tiles : 1));
    memcpy(cursor, offsets, sizeof(int64_t) * (size_t)tiles);
    for (int k = 0; k < ordered; ++k) {
        int id = order[k];
        const int* r = rects + 4 * id;
        for (int ty = r[1]; ty < r[3]; ++ty) {
            for (int tx = r[0]; tx < r[2]; ++tx) {
                ids[cursor[ty * cam->tile_count_x + tx]++] = id;
            }
        }
    }
    free(cursor);
    *offsets_out = offsets;
    return ids;
}

int oracle_render_f64(const oracle_camera* cam, int count, const double* means, const double* conics,
    const double* colors, const double* depths, const int* rects, const uint8_t* visible, const float* background,
    double* image, uint8_t* ambiguous) {
    int ordered = 0;
    int* order = depth_order(count, depths, visible, &ordered);
    int64_t* offsets = NULL;
    int* ids = tile_buckets(cam, order, ordered, rects, &offsets);
    int ambiguous_count = 0;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 4) reduction(+ : ambiguous_count)
    for (int y = 0; y < cam->height; ++y) {
        for (int x = 0; x < cam->width; ++x) {
            int tile = (y / 16) * cam->tile_count_x + x / 16;
            double T = 1.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
            int flag = 0;
            for (int64_t r = offsets[tile]; r < offsets[tile + 1]; ++r) {
                int id = ids[r];
                double dx = means[2 * id] - x, dy = means[2 * id + 1] - y;
                const double* q = conics + 4 * id;
                double t1 = q[0] * dx * dx, t2 = q[2] * dy * dy, t3 = q[1] * dx * dy;
                double power = -0.5 * (t1 + t2) - t3;
                if (fabs(power) <= 1e-5 * (fabs(t1) + fabs(t2) + 2.0 * fabs(t3)) + 1e-30) {
                    flag = 1;
                }
                double alpha = fmin(0.99, q[3] * exp(power));
                if (power > 0.0) {
                    continue;
                }
                if (near_threshold(alpha, 1.0 / 255.0, 1e-4)) {
                    flag = 1;
                }
                if (alpha < 1.0 / 255.0) {
                    continue;
                }
                double next = T * (1.0 - alpha);
                if (near_threshold(next, 1e-4, 1e-4)) {
                    flag = 1;
                }
                if (next < 1e-4) {
                    break;
                }
                double w = T * alpha;
                c0 += w * colors[3 * id];
                c1 += w * colors[3 * id + 1];
                c2 += w * colors[3 * id + 2];
                T = next;
            }
            double* pixel = image + ((int64_t)y * cam->width + x) * 3;
            pixel[0] = c0 + T * background[0];
            pixel[1] = c1 + T * background[1];
            pixel[2] = c2 + T * background[2];
            if (ambiguous) {
                ambiguous[(int64_t)y * cam->width + x] = (uint8_t)flag;
            }
            ambiguous_count += flag;
        }
    }
    free(ids);
    free(offsets);
    free(order);
    return ambiguous_count;
}

/* ---- loss and image gradients --------------------------------------------------------------- */

static void gaussian_window(double w[11]) {
    double sum = 0.0;
    for (int i = 0; i < 11; ++i) {
        double x = i - 5;
        w[i] = exp(-x * x / (2.0 * 1.5 * 1.5));
        sum += w[i];
    }
    for (int i = 0; i < 11; ++i) {
        w[i] /= sum;
    }
}

/* Separable 11-tap filter of one channel with zero padding (sepFilter2D, BORDER_CONSTANT). */
static void filter_f64(const double* input, double* output, int width, int height, const double w[11]) {
    double* temp = (double*)malloc(sizeof(double) * (size_t)width * height);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            double s = 0.0;
            for (int k = 0; k < 11; ++k) {
                int xx = x + k - 5;
                if (xx >= 0 && xx < width) {
                    s += w[k] * input[(int64_t)y * width + xx];
                }
            }
            temp[(int64_t)y * width + x] = s;
        }
    }
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int y = 0; y < height; ++y) {
        for (int x = 0; x < width; ++x) {
            double s = 0.0;
            for (int k = 0; k < 11; ++k) {
                int yy = y + k - 5;
                if (yy >= 0 && yy < height) {
                    s += w[k] * temp[(int64_t)yy * width + x];
                }
            }
            output[(int64_t)y * width + x] = s;
        }
    }
    free(temp);
}

double oracle_loss_f64(const double* image, const uint8_t* target, int width, int height) {
    double w[11];
    gaussian_window(w);
    size_t pixels = (size_t)width * height;
    double* a = (double*)malloc(sizeof(double) * pixels);
    double* b = (double*)malloc(sizeof(double) * pixels);
    double* buf[5];
    for (int k = 0; k < 5; ++k) {
        buf[k] = (double*)malloc(sizeof(double) * pixels);
    }
    double* prod = (double*)malloc(sizeof(double) * pixels);
    double l1 = 0.0, ssim = 0.0;
    for (int ch = 0; ch < 3; ++ch) {
        for (size_t p = 0; p < pixels; ++p) {
            a[p] = image[p * 3 + ch];
            b[p] = target[p * 3 + ch] / 255.0;
            l1 += fabs(a[p] - b[p]);
        }
        filter_f64(a, buf[0], width, height, w);
        filter_f64(b, buf[1], width, height, w);
        for (size_t p = 0; p < pixels; ++p) prod[p] = a[p] * a[p];
        filter_f64(prod, buf[2], width, height, w);
        for (size_t p = 0; p < pixels; ++p) prod[p] = b[p] * b[p];
        filter_f64(prod, buf[3], width, height, w);
        for (size_t p = 0; p < pixels; ++p) prod[p] = a[p] * b[p];
        filter_f64(prod, buf[4], width, height, w);
        for (size_t p = 0; p < pixels; ++p) {
            double ma = buf[0][p], mb = buf[1][p];
            double va = buf[2][p] - ma * ma, vb = buf[3][p] - mb * mb, cov = buf[4][p] - ma * mb;
            ssim += (2 * ma * mb + 0.0001) * (2 * cov + 0.0009) / ((ma * ma + mb * mb + 0.0001) * (va + vb + 0.0009));
        }
    }
    double values = 3.0 * (double)pixels;
    free(a);
    free(b);
    free(prod);
    for (int k = 0; k < 5; ++k) {
        free(buf[k]);
    }
    return 0.8 * l1 / values + 0.2 * (1.0 - ssim / values);
}

void oracle_image_gradients_f64(const double* image, const uint8_t* target, int width, int height, double* gradients) {
    double w[11];
    gaussian_window(w);
    size_t pixels = (size_t)width * height;
    double* a = (double*)malloc(sizeof(double) * pixels);
    double* b = (double*)malloc(sizeof(double) * pixels);
    double* m[5];
    double* k3[3];
    double* f3[3];
    for (int k = 0; k < 5; ++k) m[k] = (double*)malloc(sizeof(double) * pixels);
    for (int k = 0; k < 3; ++k) {
        k3[k] = (double*)malloc(sizeof(double) * pixels);
        f3[k] = (double*)malloc(sizeof(double) * pixels);
    }
    double* prod = (double*)malloc(sizeof(double) * pixels);
    double normalization = 1.0 / (3.0 * (double)width * (double)height);
    const int64_t count = (int64_t)pixels;
    for (int ch = 0; ch < 3; ++ch) {
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) {
            a[p] = image[p * 3 + ch];
            b[p] = target[p * 3 + ch] / 255.0;
        }
        filter_f64(a, m[0], width, height, w);
        filter_f64(b, m[1], width, height, w);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) prod[p] = a[p] * a[p];
        filter_f64(prod, m[2], width, height, w);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) prod[p] = b[p] * b[p];
        filter_f64(prod, m[3], width, height, w);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) prod[p] = a[p] * b[p];
        filter_f64(prod, m[4], width, height, w);
        /* solver.cu:112-123: derivative coefficients per window position. */
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) {
            double mu_a = m[0][p], mu_b = m[1][p];
            double inverse_d = 1.0 / (mu_a * mu_a + mu_b * mu_b + 0.0001);
            double inverse_e = 1.0 / (m[2][p] - mu_a * mu_a + m[3][p] - mu_b * mu_b + 0.0009);
            double luminance = (2 * mu_a * mu_b + 0.0001) * inverse_d;
            double contrast = (2 * (m[4][p] - mu_a * mu_b) + 0.0009) * inverse_e;
            double mean_derivative = 2 * (mu_b - mu_a * luminance) * inverse_d * contrast;
            double cb = 2 * luminance * inverse_e;
            double cc = -luminance * contrast * inverse_e;
            k3[0][p] = mean_derivative - mu_b * cb - 2 * mu_a * cc;
            k3[1][p] = cb;
            k3[2][p] = cc;
        }
        /* The window is symmetric, so the adjoint filter is the same zero-padded filter. */
        for (int k = 0; k < 3; ++k) {
            filter_f64(k3[k], f3[k], width, height, w);
        }
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t p = 0; p < count; ++p) {
            float render = (float)image[p * 3 + ch];
            float reference = (float)target[p * 3 + ch] * (1.0f / 255.0f);
            float residual = render - reference;
            double l1_gradient = (double)(residual > 0.0f) - (double)(residual < 0.0f);
            double ssim_gradient = -(f3[0][p] + b[p] * f3[1][p] + 2 * a[p] * f3[2][p]);
            gradients[p * 3 + ch] = (0.8 * l1_gradient + 0.2 * ssim_gradient) * normalization;
        }
    }
    free(a);
    free(b);
    free(prod);
    for (int k = 0; k < 5; ++k) free(m[k]);
    for (int k = 0; k < 3; ++k) {
        free(k3[k]);
        free(f3[k]);
    }
}

/* ---- splat backward (solver.cu:325-466), FP64 ------------------------------------------- */

/* Pixel rows [y_begin, y_end) of the replay, accumulated into screen / magnitude / tainted. */
static void backward_rows(const oracle_camera* cam, int count, const double* means, const double* conics,
    const double* colors, const int64_t* offsets, const int* ids, const double* image, const double* image_gradients,
    const uint8_t* pixel_flags, int y_begin, int y_end, double* screen, double* magnitude, uint8_t* tainted) {
    const int64_t n = count;
    int* touched = (int*)malloc(sizeof(int) * 4096);
    for (int y = y_begin; y < y_end; ++y) {
        for (int x = 0; x < cam->width; ++x) {
            int tile = (y / 16) * cam->tile_count_x + x / 16;
            const double* final_color = image + ((int64_t)y * cam->width + x) * 3;
            const double* g = image_gradients + ((int64_t)y * cam->width + x) * 3;
            double T = 1.0, prefix[3] = {0.0, 0.0, 0.0};
            /* Relative condition number of T = prod(1 - alpha_i) with respect to relative errors in
             * the alphas: sum alpha_i / (1 - alpha_i). FP32 evaluations of T (and of everything
             * built on it) carry errors proportional to it. */
            double kappa = 0.0;
            int flag = pixel_flags ? pixel_flags[(int64_t)y * cam->width + x] != 0 : 0;
            int touched_count = 0;
            for (int64_t r = offsets[tile]; r < offsets[tile + 1]; ++r) {
                int id = ids[r];
                double dx = means[2 * id] - x, dy = means[2 * id + 1] - y;
                const double* q = conics + 4 * id;
                double t1 = q[0] * dx * dx, t2 = q[2] * dy * dy, t3 = q[1] * dx * dy;
                double power = -0.5 * (t1 + t2) - t3;
                if (fabs(power) <= 1e-5 * (fabs(t1) + fabs(t2) + 2.0 * fabs(t3)) + 1e-30) {
                    flag = 1;
                }
                if (power > 0.0) {
                    continue;
                }
                double raw_alpha = q[3] * exp(power);
                double alpha = fmin(0.99, raw_alpha);
                if (near_threshold(alpha, 1.0 / 255.0, 1e-4) || near_threshold(raw_alpha, 0.99, 1e-4)) {
                    flag = 1;
                }
                if (alpha < 1.0 / 255.0) {
                    continue;
                }
                double next = T * (1.0 - alpha);
                if (near_threshold(next, 1e-4, 1e-4)) {
                    flag = 1;
                }
                if (next < 1e-4) {
                    break;
                }
                if (touched_count < 4096) {
                    touched[touched_count++] = id;
                }
                const double* col = colors + 3 * id;
                double weight = alpha * T;
                double alpha_gradient = 0.0;
                /* Error-bound magnitude of alpha_gradient: the sum of the absolute values of the
                 * summands prefix + T c - C, which cancel when T is small. An FP32 evaluation of
                 * the same expression is accurate relative to this, not to |alpha_gradient|. */
                double alpha_magnitude = 0.0;
                double term;
                for (int c = 0; c < 3; ++c) {
                    alpha_gradient += g[c] * (prefix[c] + T * col[c] - final_color[c]) / (1.0 - alpha);
                    alpha_magnitude +=
                        fabs(g[c]) * (fabs(prefix[c]) + fabs(T * col[c]) + fabs(final_color[c])) / (1.0 - alpha);
                    term = weight * g[c];
                    screen[(6 + c) * n + id] += term;
                    magnitude[(6 + c) * n + id] += fabs(term) * (1.0 + kappa);
                }
                if (raw_alpha < 0.99) {
                    double common = alpha_gradient * raw_alpha;
                    double common_magnitude = alpha_magnitude * raw_alpha * (1.0 + kappa + alpha / (1.0 - alpha));
                    double qx = q[0] * dx + q[1] * dy;
                    double qy = q[1] * dx + q[2] * dy;
                    double terms[6] = {-common * qx, -common * qy, 0.5 * common * qx * qx, common * qx * qy,
                        0.5 * common * qy * qy, common * (1.0 - q[3])};
                    double bounds[6] = {fabs(qx), fabs(qy), 0.5 * qx * qx, fabs(qx * qy), 0.5 * qy * qy,
                        fabs(1.0 - q[3])};
                    for (int k = 0; k < 6; ++k) {
                        screen[k * n + id] += terms[k];
                        magnitude[k * n + id] += common_magnitude * bounds[k];
                    }
                    screen[9 * n + id] += fabs(terms[0]);
                    screen[10 * n + id] += fabs(terms[1]);
                    magnitude[9 * n + id] += common_magnitude * bounds[0];
                    magnitude[10 * n + id] += common_magnitude * bounds[1];
                }
                for (int c = 0; c < 3; ++c) {
                    prefix[c] += col[c] * weight;
                }
                T = next;
                kappa += alpha / (1.0 - alpha);
            }
            if (flag) {
                for (int k = 0; k < touched_count; ++k) {
                    tainted[touched[k]] = 1;
                }
            }
        }
    }
    free(touched);
}

void oracle_backward_f64(const oracle_camera* cam, int count, const double* means, const double* conics,
    const double* colors, const double* depths, const int* rects, const uint8_t* visible, const double* image,
    const double* image_gradients, const uint8_t* pixel_flags, double* screen, double* magnitude,
    uint8_t* tainted) {
    const size_t n = (size_t)count;
    memset(screen, 0, sizeof(double) * 11 * n);
    memset(magnitude, 0, sizeof(double) * 11 * n);
    memset(tainted, 0, n);
    int ordered = 0;
    int* order = depth_order(count, depths, visible, &ordered);
    int64_t* offsets = NULL;
    int* ids = tile_buckets(cam, order, ordered, rects, &offsets);
    int threads = g_threads < cam->height ? g_threads : (cam->height > 0 ? cam->height : 1);
    if (threads <= 1) {
        backward_rows(cam, count, means, conics, colors, offsets, ids, image, image_gradients, pixel_flags, 0,
            cam->height, screen, magnitude, tainted);
    }
    else {
        /* Thread t replays a contiguous band of rows into its own buffers (thread 0 into the
         * outputs); the bands are then added in thread order. */
        double* partial_screen = (double*)calloc(11 * n * (size_t)(threads - 1), sizeof(double));
        double* partial_magnitude = (double*)calloc(11 * n * (size_t)(threads - 1), sizeof(double));
        uint8_t* partial_tainted = (uint8_t*)calloc(n * (size_t)(threads - 1), 1);
#pragma omp parallel for num_threads(threads) schedule(static, 1)
        for (int t = 0; t < threads; ++t) {
            int y_begin = (int)((int64_t)cam->height * t / threads);
            int y_end = (int)((int64_t)cam->height * (t + 1) / threads);
            double* s_out = t == 0 ? screen : partial_screen + 11 * n * (size_t)(t - 1);
            double* m_out = t == 0 ? magnitude : partial_magnitude + 11 * n * (size_t)(t - 1);
            uint8_t* f_out = t == 0 ? tainted : partial_tainted + n * (size_t)(t - 1);
            backward_rows(cam, count, means, conics, colors, offsets, ids, image, image_gradients, pixel_flags,
                y_begin, y_end, s_out, m_out, f_out);
        }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < (int64_t)(11 * n); ++i) {
            for (int t = 1; t < threads; ++t) {
                screen[i] += partial_screen[11 * n * (size_t)(t - 1) + (size_t)i];
                magnitude[i] += partial_magnitude[11 * n * (size_t)(t - 1) + (size_t)i];
            }
        }
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) {
            for (int t = 1; t < threads; ++t) {
                tainted[i] |= partial_tainted[n * (size_t)(t - 1) + (size_t)i];
            }
        }
        free(partial_screen);
        free(partial_magnitude);
        free(partial_tainted);
    }
    free(ids);
    free(offsets);
    free(order);
}

/* ---- chain rule (solver.cu:468-623), FP64 ------------------------------------------------ */

void oracle_map_f64(const float* params, int count, const oracle_camera* cam, int sh_degree, const double* screen,
    const uint8_t* observed, double* gradients, double* statistics, double* absolute_statistics) {
    sh_degree = sh_degree < 0 ? 0 : (sh_degree > 3 ? 3 : sh_degree);
    const int s = (sh_degree + 1) * (sh_degree + 1);
    const int64_t n = count;
    double Rc[9], tc[3];
    for (int k = 0; k < 9; ++k) Rc[k] = cam->R_c_w[k];
    for (int k = 0; k < 3; ++k) tc[k] = cam->t_c_w[k];
    memset(gradients, 0, sizeof(double) * PARAMS * (size_t)count);
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int id = 0; id < count; ++id) {
        double in[11];
        for (int k = 0; k < 11; ++k) in[k] = screen[k * n + id];
        double sx = in[0] * (0.5 * cam->width), sy = in[1] * (0.5 * cam->height);
        double ax = in[9] * (0.5 * cam->width), ay = in[10] * (0.5 * cam->height);
        statistics[id] = sqrt(sx * sx + sy * sy);
        statistics[n + id] = observed[id] ? 1.0 : 0.0;
        absolute_statistics[id] = sqrt(ax * ax + ay * ay);
        absolute_statistics[n + id] = in[10];
        int active = 0;
        for (int k = 0; k < 9; ++k) active = active || in[k] != 0.0;
        if (!active) continue;
        double omega[3], world[3], rotation[9], position[3], CR[9], s2[3], cov[9];
        for (int k = 0; k < 3; ++k) {
            omega[k] = params[(ROTATION_INDEX + k) * n + id];
            world[k] = params[k * n + id];
        }
        rotation_f64(omega[0], omega[1], omega[2], rotation);
        for (int r = 0; r < 3; ++r) {
            position[r] = tc[r];
            for (int c = 0; c < 3; ++c) {
                position[r] += Rc[r * 3 + c] * world[c];
                CR[r * 3 + c] = Rc[r * 3] * rotation[c] + Rc[r * 3 + 1] * rotation[3 + c] + Rc[r * 3 + 2] * rotation[6 + c];
            }
            double sc = exp((double)params[(LOG_SCALE_INDEX + r) * n + id]);
            s2[r] = sc * sc;
        }
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) {
                double sum = 0.0;
                for (int j = 0; j < 3; ++j) sum += CR[r * 3 + j] * s2[j] * CR[c * 3 + j];
                cov[r * 3 + c] = sum;
            }
        }
        double iz = 1.0 / position[2];
        double nx = position[0] * iz, ny = position[1] * iz;
        double lx = 0.65 * cam->width / cam->fx, ly = 0.65 * cam->height / cam->fy;
        double jx = fmin(fmax(nx, -lx), lx), jy = fmin(fmax(ny, -ly), ly);
        double P[6] = {cam->fx * iz, 0.0, -cam->fx * jx * iz, 0.0, cam->fy * iz, -cam->fy * jy * iz};
        double product[6], LP[6];
        for (int r = 0; r < 2; ++r) {
            for (int c = 0; c < 3; ++c) {
                double p = 0.0, l = 0.0;
                for (int j = 0; j < 3; ++j) {
                    p += P[r * 3 + j] * cov[j * 3 + c];
                    l += P[r * 3 + j] * CR[j * 3 + c];
                }
                product[r * 3 + c] = p;
                LP[r * 3 + c] = l;
            }
        }
        /* ComputeEquationSh<D> (solver.cu:223-323) */
        double dir[3];
        for (int a = 0; a < 3; ++a) {
            dir[a] = world[a];
            for (int b = 0; b < 3; ++b) dir[a] += Rc[b * 3 + a] * tc[b];
        }
        double il = 1.0 / sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        double x = dir[0] * il, y = dir[1] * il, z = dir[2] * il;
        double basis[16], d[48];
        sh_basis_f64(x, y, z, basis);
        memset(d, 0, sizeof(d));
        {
            double C1 = 0.4886025119029199;
            d[4] = -C1; d[8] = C1; d[9] = -C1;
            double c = 1.0925484305920792, dd = 0.31539156525252005, e = 0.5462742152960396;
            d[12] = c * y; d[13] = c * x; d[16] = -c * z; d[17] = -c * y;
            d[18] = -2 * dd * x; d[19] = -2 * dd * y; d[20] = 4 * dd * z;
            d[21] = -c * z; d[23] = -c * x; d[24] = 2 * e * x; d[25] = -2 * e * y;
            double a3 = -0.5900435899266435, b3 = 2.890611442640554, c3 = -0.4570457994644658;
            double d3 = 0.3731763325901154, e3 = 1.445305721320277;
            d[27] = 6 * a3 * x * y; d[28] = 3 * a3 * (x * x - y * y);
            d[30] = b3 * y * z; d[31] = b3 * x * z; d[32] = b3 * x * y;
            d[33] = -2 * c3 * x * y; d[34] = c3 * (4 * z * z - x * x - 3 * y * y); d[35] = 8 * c3 * y * z;
            d[36] = -6 * d3 * x * z; d[37] = -6 * d3 * y * z; d[38] = d3 * (6 * z * z - 3 * x * x - 3 * y * y);
            d[39] = c3 * (4 * z * z - 3 * x * x - y * y); d[40] = -2 * c3 * x * y; d[41] = 8 * c3 * x * z;
            d[42] = 2 * e3 * x * z; d[43] = -2 * e3 * y * z; d[44] = e3 * (x * x - y * y);
            d[45] = 3 * a3 * (x * x - y * y); d[46] = -6 * a3 * x * y;
        }
        double color[3], cd[9];
        for (int ch = 0; ch < 3; ++ch) {
            double value = 0.0, der[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < s; ++j) {
                double coefficient = params[(SH_INDEX + ch * 16 + j) * n + id];
                value += coefficient * basis[j];
                for (int a = 0; a < 3; ++a) der[a] += coefficient * d[j * 3 + a];
            }
            color[ch] = value + 0.5;
            double radial = der[0] * x + der[1] * y + der[2] * z;
            cd[ch * 3 + 0] = (der[0] - radial * x) * il;
            cd[ch * 3 + 1] = (der[1] - radial * y) * il;
            cd[ch * 3 + 2] = (der[2] - radial * z) * il;
        }
        double rg[3];
        for (int axis = 0; axis < 3; ++axis) {
            double dx = Rc[axis], dy = Rc[3 + axis], dz = Rc[6 + axis];
            double du = (dx - nx * dz) * iz, dv = (dy - ny * dz) * iz;
            double djx = (nx >= -lx && nx <= lx) ? du : 0.0;
            double djy = (ny >= -ly && ny <= ly) ? dv : 0.0;
            double delta[6] = {-cam->fx * dz * iz * iz, 0.0, -cam->fx * (djx - jx * dz * iz) * iz, 0.0,
                -cam->fy * dz * iz * iz, -cam->fy * (djy - jy * dz * iz) * iz};
            double daa = 0.0, dab = 0.0, ddd = 0.0;
            for (int j = 0; j < 3; ++j) {
                daa += 2 * delta[j] * product[j];
                dab += delta[j] * product[3 + j] + delta[3 + j] * product[j];
                ddd += 2 * delta[3 + j] * product[3 + j];
            }
            double center = in[0] * cam->fx * du + in[1] * cam->fy * dv + in[2] * daa + in[3] * dab + in[4] * ddd;
            for (int c = 0; c < 3; ++c) {
                if (color[c] >= 0) center += in[6 + c] * cd[c * 3 + axis];
            }
            gradients[axis * n + id] = center;
            double a = LP[axis], b = LP[3 + axis];
            gradients[(LOG_SCALE_INDEX + axis) * n + id] = 2 * s2[axis] * (in[2] * a * a + in[3] * a * b + in[4] * b * b);
            int j = (axis + 1) % 3, k = (axis + 2) % 3;
            rg[axis] = (s2[j] - s2[k]) * (2 * in[2] * LP[j] * LP[k] + in[3] * (LP[j] * LP[3 + k] + LP[k] * LP[3 + j]) +
                2 * in[4] * LP[3 + j] * LP[3 + k]);
        }
        double theta2 = omega[0] * omega[0] + omega[1] * omega[1] + omega[2] * omega[2];
        double sa = 0.5, sb = 1.0 / 6.0;
        if (theta2 > 1e-4) {
            double theta = sqrt(theta2);
            double hs = sin(0.5 * theta) / (0.5 * theta);
            sa = 0.5 * hs * hs;
            sb = (theta - sin(theta)) / (theta2 * theta);
        }
        else {
            sa -= theta2 / 24.0;
            sb -= theta2 / 120.0;
        }
        double cross[3];
        for (int axis = 0; axis < 3; ++axis) {
            int j = (axis + 1) % 3, k = (axis + 2) % 3;
            cross[axis] = omega[j] * rg[k] - omega[k] * rg[j];
        }
        for (int axis = 0; axis < 3; ++axis) {
            int j = (axis + 1) % 3, k = (axis + 2) % 3;
            gradients[(ROTATION_INDEX + axis) * n + id] = rg[axis] + sa * cross[axis] + sb * (omega[j] * cross[k] - omega[k] * cross[j]);
        }
        gradients[OPACITY_INDEX * n + id] = in[5];
        for (int c = 0; c < 3; ++c) {
            for (int k = 0; k < s; ++k) {
                gradients[(SH_INDEX + c * 16 + k) * n + id] = color[c] >= 0 ? in[6 + c] * basis[k] : 0.0;
            }
        }
    }
}

/* ---- Adam and history remap (solver.cu:625-689), FP32 mirror ----------------------------- */

static float learning_rate_for_row(int row, float position_learning_rate) {
    if (row < LOG_SCALE_INDEX) return position_learning_rate;
    if (row < ROTATION_INDEX) return 0.005f;
    if (row < OPACITY_INDEX) return 0.001f;
    if (row < SH_INDEX) return 0.025f;
    return (row - SH_INDEX) % 16 == 0 ? 0.0025f : 0.0025f / 20.0f;
}

void oracle_adam_f32(int count, const float* gradients, float* params, float* first, float* second,
    float position_learning_rate, float first_correction, float second_correction) {
    for (int row = 0; row < PARAMS; ++row) {
        float lr = learning_rate_for_row(row, position_learning_rate);
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int id = 0; id < count; ++id) {
            int64_t index = (int64_t)row * count + id;
            float g = gradients[index];
            float m = 0.9f * first[index] + 0.1f * g;
            float v = 0.999f * second[index] + 0.001f * g * g;
            first[index] = m;
            second[index] = v;
            params[index] -= lr * (m * first_correction) / (sqrtf(v * second_correction) + 1e-15f);
        }
    }
}

void oracle_remap_f32(const int* parent_ids, int previous_count, int count, const float* previous_first,
    const float* previous_second, float* first, float* second) {
    for (int row = 0; row < PARAMS; ++row) {
        for (int id = 0; id < count; ++id) {
            int parent = parent_ids ? parent_ids[id] : -1;
            int64_t index = (int64_t)row * count + id;
            if (parent >= 0 && parent < previous_count) {
                first[index] = previous_first[(int64_t)row * previous_count + parent];
                second[index] = previous_second[(int64_t)row * previous_count + parent];
            }
            else {
                first[index] = 0.0f;
                second[index] = 0.0f;
            }
        }
    }
}

/* ---- density control (density_controller.cu:60-278), FP32 mirror ------------------------- */

/* proposal == NULL: the reference rule. Otherwise logits[N][3] (keep, clone, split) choose the
 * densification by the largest logit (ties and NaN favour keep, then clone); pruning is unchanged.
 * The proposal form has no reference counterpart. */
static void density_count(const float* params, int count, const float* sums, const float* abs_sums,
    const uint32_t* counts, float gradient_threshold, float split_threshold, float min_opacity, float max_world_size,
    int prune_large, float absolute_gradient_threshold, const float* proposal, int32_t* decisions, int32_t* offsets) {
    const int64_t n = count;
    for (int id = 0; id < count; ++id) {
        float opacity = 1.0f / (1.0f + oracle_expf(-params[OPACITY_INDEX * n + id]));
        uint32_t observations = counts ? counts[id] : 0u;
        float gradient = observations ? sums[id] / (float)observations : 0.0f;
        if (gradient != gradient) gradient = 0.0f;
        float scale[3];
        for (int a = 0; a < 3; ++a) scale[a] = oracle_expf(params[(LOG_SCALE_INDEX + a) * n + id]);
        float max_scale = fmaxf(scale[0], fmaxf(scale[1], scale[2]));
        if (scale[0] != scale[0] || scale[1] != scale[1] || scale[2] != scale[2]) max_scale = NAN;
        int clone, split;
        if (proposal) {
            float k = proposal[3 * (int64_t)id], c = proposal[3 * (int64_t)id + 1], s = proposal[3 * (int64_t)id + 2];
            clone = c > k && c >= s;
            split = s > k && s > c;
        } else {
            int selected = gradient >= gradient_threshold;
            clone = selected && max_scale <= split_threshold;
            float absolute_gradient = observations ? abs_sums[id] / (float)observations : 0.0f;
            int split_selected =
                absolute_gradient_threshold > 0.0f ? absolute_gradient >= absolute_gradient_threshold : selected;
            split = split_selected && max_scale > split_threshold;
        }
        float remaining = max_scale;
        if (split && prune_large) {
            float x = oracle_expf(oracle_logf(scale[0] / 1.6f));
            float y = oracle_expf(oracle_logf(scale[1] / 1.6f));
            float z = oracle_expf(oracle_logf(scale[2] / 1.6f));
            remaining = fmaxf(x, fmaxf(y, z));
            if (x != x || y != y || z != z) remaining = NAN;
        }
        int remove = opacity < min_opacity || (prune_large && remaining > max_world_size);
        decisions[4 * id + 0] = !split && !remove;
        decisions[4 * id + 1] = clone && !remove;
        decisions[4 * id + 2] = split;
        decisions[4 * id + 3] = split && !remove;
    }
    for (int k = 0; k < 4; ++k) decisions[4 * count + k] = 0;
    int running[4] = {0, 0, 0, 0};
    for (int id = 0; id <= count; ++id) {
        for (int k = 0; k < 4; ++k) {
            offsets[4 * id + k] = running[k];
            running[k] += decisions[4 * id + k];
        }
    }
}

void oracle_density_count_f32(const float* params, int count, const float* sums, const float* abs_sums,
    const uint32_t* counts, float gradient_threshold, float split_threshold, float min_opacity, float max_world_size,
    int prune_large, float absolute_gradient_threshold, int32_t* decisions, int32_t* offsets) {
    density_count(params, count, sums, abs_sums, counts, gradient_threshold, split_threshold, min_opacity,
        max_world_size, prune_large, absolute_gradient_threshold, NULL, decisions, offsets);
}

void oracle_density_count_proposed_f32(const float* params, int count, const float* proposal, float split_threshold,
    float min_opacity, float max_world_size, int prune_large, int32_t* decisions, int32_t* offsets) {
    density_count(params, count, NULL, NULL, NULL, 0.0f, split_threshold, min_opacity, max_world_size, prune_large,
        0.0f, proposal, decisions, offsets);
}

void oracle_philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = (uint32_t)p1;
        c[2] = n2;
        c[3] = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

float oracle_jitter_normal(uint64_t seed, uint64_t counter, int64_t index) {
    uint64_t block = (uint64_t)index >> 2;
    uint32_t c[4] = {(uint32_t)block, (uint32_t)(block >> 32), (uint32_t)counter, (uint32_t)(counter >> 32)};
    oracle_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    int component = (int)(index & 3);
    int pair = component >> 1;
    float u1 = (float)(c[2 * pair] >> 8) * 0x1p-24f + 0x1p-25f;
    float u2 = (float)(c[2 * pair + 1] >> 8) * 0x1p-24f + 0x1p-25f;
    float radius = sqrtf(-2.0f * oracle_logf(u1));
    float s, co;
    oracle_sincosf(6.28318530717958647692f * u2, &s, &co);
    return (component & 1) ? radius * s : radius * co;
}

void oracle_density_write_f32(const float* params, int count, const int32_t* decisions, const int32_t* offsets,
    uint64_t seed, uint64_t counter, float* out, int out_count, int32_t* history_parent_ids, int32_t* parent_ids,
    float* child_flags) {
    const int64_t n = count, m = out_count;
    const int32_t* totals = offsets + 4 * count;
    for (int id = 0; id < count; ++id) {
        const int32_t* d = decisions + 4 * id;
        const int32_t* o = offsets + 4 * id;
        int split = d[3] != 0;
        if (!d[0] && !d[1] && !split) continue;
        float scale[3] = {0.0f, 0.0f, 0.0f}, rotation[9] = {0};
        if (split) {
            for (int a = 0; a < 3; ++a) scale[a] = oracle_expf(params[(LOG_SCALE_INDEX + a) * n + id]);
            float x = params[(ROTATION_INDEX + 0) * n + id];
            float y = params[(ROTATION_INDEX + 1) * n + id];
            float z = params[(ROTATION_INDEX + 2) * n + id];
            float angle = sqrtf(x * x + y * y + z * z);
            float a = 1.0f, b = 0.5f;
            if (angle > 1e-4f) {
                a = oracle_sinf(angle) / angle;
                float half_sinc = oracle_sinf(angle * 0.5f) / (angle * 0.5f);
                b = half_sinc * half_sinc * 0.5f;
            }
            else {
                float square = angle * angle;
                a -= square / 6.0f;
                b -= square / 24.0f;
            }
            rotation[0] = 1.0f - b * (y * y + z * z);
            rotation[1] = b * x * y - a * z;
            rotation[2] = b * x * z + a * y;
            rotation[3] = b * x * y + a * z;
            rotation[4] = 1.0f - b * (x * x + z * z);
            rotation[5] = b * y * z - a * x;
            rotation[6] = b * x * z - a * y;
            rotation[7] = b * y * z + a * x;
            rotation[8] = 1.0f - b * (x * x + y * y);
        }
        int children = split ? 2 : d[0] + d[1];
        for (int child = 0; child < children; ++child) {
            int destination = split ? totals[0] + totals[1] + child * totals[3] + o[3]
                                    : (child == 0 && d[0] ? o[0] : totals[0] + o[1]);
            float normals[3] = {0.0f, 0.0f, 0.0f};
            if (split) {
                for (int a = 0; a < 3; ++a) {
                    normals[a] = oracle_jitter_normal(seed, counter, ((int64_t)child * totals[2] + o[2]) * 3 + a);
                }
            }
            for (int p = 0; p < PARAMS; ++p) {
                float value = params[p * n + id];
                if (split && p < 3) {
                    float delta = 0.0f;
                    for (int a = 0; a < 3; ++a) {
                        float local = normals[a] * scale[a];
                        delta += rotation[p * 3 + a] * local;
                    }
                    value += delta;
                }
                else if (split && p >= LOG_SCALE_INDEX && p < LOG_SCALE_INDEX + 3) {
                    value = oracle_logf(scale[p - LOG_SCALE_INDEX] / 1.6f);
                }
                out[p * m + destination] = value;
            }
            if (history_parent_ids) history_parent_ids[destination] = !split && child == 0 && d[0] ? id : -1;
            if (parent_ids) parent_ids[destination] = id;
            if (child_flags) child_flags[destination] = (d[1] || d[2]) ? 1.0f : 0.0f;
        }
    }
}

/* ---- KNN (knn.cu:180-255) by exhaustive search ------------------------------------------ */

void oracle_knn3(const float* points, int n, int32_t* neighbors, float* mean_dist2) {
    if (n < 2) {
        for (int i = 0; i < n; ++i) {
            mean_dist2[i] = 0.0f;
            if (neighbors) neighbors[3 * i] = neighbors[3 * i + 1] = neighbors[3 * i + 2] = -1;
        }
        return;
    }
    const float inf = bits_to_float(0x7f800000u);
    for (int i = 0; i < n; ++i) {
        float bd[3] = {inf, inf, inf};
        int bi[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
        const float* p = points + 3 * (int64_t)i;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const float* q = points + 3 * (int64_t)j;
            float dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
            float d = dx * dx + dy * dy + dz * dz;
            /* insert (d, j) into the (distance, index)-ordered best three */
            int pos = 3;
            while (pos > 0 && (d < bd[pos - 1] || (d == bd[pos - 1] && j < bi[pos - 1]))) --pos;
            if (pos < 3) {
                for (int k = 2; k > pos; --k) {
                    bd[k] = bd[k - 1];
                    bi[k] = bi[k - 1];
                }
                bd[pos] = d;
                bi[pos] = j;
            }
        }
        int others = n - 1;
        float result;
        if (others >= 3) {
            result = (bd[0] + bd[1] + bd[2]) / 3.0f;
        }
        else {
            float sum = bd[0];
            if (others == 2) sum += bd[1];
            result = sum / (float)others;
        }
        mean_dist2[i] = result;
        if (neighbors) {
            for (int k = 0; k < 3; ++k) neighbors[3 * i + k] = k < others ? bi[k] : -1;
        }
    }
}
