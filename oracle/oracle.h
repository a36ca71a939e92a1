/*
 * oracle.h - CPU restatement of the reference (Naive-q/GS-native) training hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The product never calls it.
 *
 * Two families of functions:
 *   *_f32  FP32 mirrors: same operation order as the device code, FP contraction off, the
 *          project-written exp/log/sincos. Used where outputs must match bit for bit
 *          (tile rectangles, sort keys, tile ranges, Adam, density masks, KNN).
 *   *_f64  FP64 restatements of the reference math (test/cpu_reference.h, renderer.cu,
 *          solver.cu). Used for tolerance checks of images, gradients and loss values, and
 *          pinned against finite differences in the tests.
 * Every function cites the reference file:line it follows (paths relative to the reference).
 */
#ifndef GSN_ORACLE_H
#define GSN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_camera {
    float fx, fy, cx, cy;
    float R_c_w[9];
    float t_c_w[3];
    float near_depth, far_depth;
    int width, height;
    int64_t pixel_offset;
    int tile_count_x, tile_count_y;
    int tile_offset;
} oracle_camera;

/* Threads of the parallel loops (OpenMP) of the FP64 iteration functions (projection, render,
 * image gradients, splat backward, chain rule) and of Adam. Default 1: sequential order. */
void oracle_set_threads(int threads);

/* Axis-angle to row-major rotation (renderer.cu:35-66) in FP32 (the device's operation order)
 * and FP64, and the 16 real SH basis values of a unit direction (renderer.cu:68-106) in FP64. */
void oracle_rotation_f32(float rx, float ry, float rz, float* R);
void oracle_rotation_f64(double rx, double ry, double rz, double* R);
void oracle_sh_basis_f64(double x, double y, double z, double* basis);

/* Project-written FP32 transcendentals (identical algorithm to the device code). */
float oracle_expf(float x);
float oracle_logf(float x);
void oracle_sincosf(float x, float* s, float* c);

/* renderer.cu:108-221 (FP32 mirror). record_counts[N+1]. */
void oracle_project_f32(const float* params, int count, const oracle_camera* camera, int sh_degree, float* means,
    float* conic_opacities, float* colors, float* depths, int* tile_rects, int* record_counts);

/* renderer.cu:465-487: exclusive scan in place; returns the total. */
int oracle_scan_i32(int* values, int count);

/* renderer.cu:223-271, 541-569: keys (tile << 32 | depth bits) enumerated in id order,
 * stable sort, tile ranges. sorted_keys may be NULL. */
void oracle_tile_lists(const oracle_camera* camera, int count, const int* tile_rects, const float* depths,
    const int* record_offsets, int record_count, int* tile_gaussian_ids, int64_t* tile_offsets,
    uint64_t* sorted_keys);

/* renderer.cu:273-338 (FP32 mirror with libm expf). ambiguous[H*W] (optional) is set to 1 for
 * pixels where a threshold decision lies within a relative margin of its threshold; returns the
 * number of such pixels. */
int oracle_render_f32(const oracle_camera* camera, const int* tile_gaussian_ids, const int64_t* tile_offsets,
    const float* means, const float* conic_opacities, const float* colors, const float* background, float* image,
    uint8_t* ambiguous);

/* FP64 projection (renderer.cu:108-221 in double, cpu_reference.h:47-79). visible[N]. */
void oracle_project_f64(const float* params, int count, const oracle_camera* camera, int sh_degree, double* means,
    double* conics, double* colors, double* depths, int* tile_rects, uint8_t* visible);

/* test/cpu_reference.h:23-118 over FP64 projections: Gaussians in stable depth order, each
 * blended into the pixels of its tile rectangle. ambiguous[H*W] optional (as above). */
int oracle_render_f64(const oracle_camera* camera, int count, const double* means, const double* conics,
    const double* colors, const double* depths, const int* tile_rects, const uint8_t* visible,
    const float* background, double* image, uint8_t* ambiguous);

/* test/cpu_reference.h:120-152: 0.8 * L1 + 0.2 * (1 - SSIM), 11-tap sigma 1.5, zero border. */
double oracle_loss_f64(const double* image, const uint8_t* target, int width, int height);

/* solver.cu:26-168 in FP64 (sign of the residual taken in FP32 as the device does). */
void oracle_image_gradients_f64(const double* image, const uint8_t* target, int width, int height,
    double* gradients);

/* solver.cu:325-466 in FP64: per-pixel front-to-back replay from the final colour. screen[11][N]
 * receives dmean2D(2), dSigma(3), dlogit, dcolor(3), |dmean2D|(2); magnitude[11][N] the sums of
 * the per-pixel terms' error-bound magnitudes, the scale for tolerances: |term| (1 + kappa) with
 * kappa = sum over the splats in front of alpha / (1 - alpha), the relative condition number of
 * T; for the geometry fields dL/dalpha's summands |g| (|prefix| + |T c| + |C|) / (1 - alpha) stand
 * in for the cancelled |dL/dalpha| and the splat's own alpha / (1 - alpha) joins kappa.
 * tainted[N] marks Gaussians that touch an ambiguous pixel; pixel_flags[H][W] (optional) adds
 * pixels the caller knows to be ambiguous (e.g. an L1 sign decided within FP32 error). */
void oracle_backward_f64(const oracle_camera* camera, int count, const double* means, const double* conics,
    const double* colors, const double* depths, const int* tile_rects, const uint8_t* visible, const double* image,
    const double* image_gradients, const uint8_t* pixel_flags, double* screen, double* magnitude,
    uint8_t* tainted);

/* solver.cu:468-490 and 492-623 in FP64: statistics[2][N], absolute_statistics[2][N] and the
 * parameter gradients[58][N] from screen[11][N] and record counts (observed = count > 0). */
void oracle_map_f64(const float* params, int count, const oracle_camera* camera, int sh_degree, const double* screen,
    const uint8_t* observed, double* gradients, double* statistics, double* absolute_statistics);

/* solver.cu:651-689 (FP32 mirror, contraction off). */
void oracle_adam_f32(int count, const float* gradients, float* params, float* first, float* second,
    float position_learning_rate, float first_correction, float second_correction);

/* solver.cu:625-649. */
void oracle_remap_f32(const int* parent_ids, int previous_count, int count, const float* previous_first,
    const float* previous_second, float* first, float* second);

/* density_controller.cu:60-116 + int4 exclusive scan; decisions/offsets int4[N+1]. */
void oracle_density_count_f32(const float* params, int count, const float* gradient_sums,
    const float* absolute_gradient_sums, const uint32_t* observation_counts, float gradient_threshold,
    float split_threshold, float min_opacity, float max_world_size, int prune_large,
    float absolute_gradient_threshold, int32_t* decisions, int32_t* offsets);

/* The same with a learned proposal: logits[N][3] (keep, clone, split) replace the gradient rule's
 * selection by their largest entry (ties and NaN favour keep, then clone); pruning unchanged.
 * No reference counterpart. */
void oracle_density_count_proposed_f32(const float* params, int count, const float* proposal,
    float split_threshold, float min_opacity, float max_world_size, int prune_large, int32_t* decisions,
    int32_t* offsets);

/* density_controller.cu:168-278 with the build's Philox4x32-10 + Box-Muller jitter. */
void oracle_density_write_f32(const float* params, int count, const int32_t* decisions, const int32_t* offsets,
    uint64_t seed, uint64_t counter, float* out, int out_count, int32_t* history_parent_ids, int32_t* parent_ids,
    float* child_flags);
float oracle_jitter_normal(uint64_t seed, uint64_t counter, int64_t index);
void oracle_philox4x32_10(uint32_t counter[4], uint32_t key0, uint32_t key1);

/* knn.cu:180-255 by exhaustive search; neighbours ordered by (squared distance, index). */
void oracle_knn3(const float* points, int n, int32_t* neighbors, float* mean_dist2);

#ifdef __cplusplus
}
#endif

#endif
