/*
 * gsn_hip.h - C ABI of the MI355X (gfx950) Gaussian-splatting training hot path.
 *
 * Every entry point takes an opaque context (one HIP stream plus the scratch memory the
 * kernels need), device pointers, POD structs and sizes, and returns an int status
 * (GSN_OK == 0). No C++ types cross this boundary. All work is enqueued on the context's
 * stream; entry points that hand a value back to the host say so and synchronise that stream.
 *
 * Layouts follow the reference (Naive-q/GS-native):
 *   params            float[58][N]  rows: 0-2 position, 3-5 log-scale, 6-8 axis-angle,
 *                                   9 opacity logit, 10+16c+j SH channel c coeff j
 *                                   (src/core/types/buffers.h:10-15)
 *   means             float[N][2]   pixel-space centres          (renderer.h:25)
 *   conic_opacities   float[N][4]   (A, B, C, sigmoid(opacity))   (renderer.h:29)
 *   colors            float[N][3]   SH colour, clamped below at 0 (renderer.h:33)
 *   depths            float[N]      camera-space z
 *   tile_rects        int[N][4]     (x0, y0, x1, y1) tile rectangle, exclusive
 *   record_offsets    int[N+1]      records per Gaussian, then their exclusive prefix
 *   tile_gaussian_ids int[R]        Gaussian id per record, tile-major then depth
 *   tile_offsets      int64[T+1]    [begin, end) of each tile's records
 *   images            float[H][W][3] interleaved RGB; targets uint8[H][W][3]
 *
 * Each declaration names the reference function it replaces (file:line in the reference).
 */
#ifndef GSN_HIP_H
#define GSN_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GSN_OK = 0,
    GSN_ERROR_INVALID_ARGUMENT = 1,
    GSN_ERROR_HIP = 2,
    GSN_ERROR_OUT_OF_MEMORY = 3,
    GSN_ERROR_CAPACITY = 4,
};

/* Byte-identical to gs_native::CameraParams (src/core/types/buffers.h:17-28). */
typedef struct gsn_camera {
    float fx, fy, cx, cy;
    float R_c_w[9];
    float t_c_w[3];
    float near_depth, far_depth;
    int width, height;
    int64_t pixel_offset;
    int tile_count_x, tile_count_y;
    int tile_offset;
} gsn_camera;

typedef struct gsn_context gsn_context;

/* ---- context, status, device plumbing -------------------------------------------------- */

/* Creates a context with its own HIP stream (default priority, blocking w.r.t. stream 0). */
int gsn_context_create(gsn_context** context);
void gsn_context_destroy(gsn_context* context);
int gsn_context_synchronize(gsn_context* context);
/* Waits for all work on the current device (every context's stream). */
int gsn_device_synchronize(void);
/* Returns the context's hipStream_t. */
void* gsn_context_stream(gsn_context* context);
/* Bytes of scratch memory currently held by the context. */
size_t gsn_context_scratch_bytes(gsn_context* context);
const char* gsn_status_string(int status);
/* Message describing the most recent failure on the calling thread ("" if none). */
const char* gsn_last_error(void);

int gsn_device_count(int* count);
int gsn_device_name(char* name, size_t capacity);
int gsn_device_malloc(void** pointer, size_t bytes);
int gsn_device_free(void* pointer);
int gsn_memcpy_to_device(void* destination, const void* source, size_t bytes);
int gsn_memcpy_to_host(void* destination, const void* source, size_t bytes);
int gsn_memcpy_device(void* destination, const void* source, size_t bytes);
int gsn_memset_device(void* destination, int value, size_t bytes);
/* Peak bytes reported by hipMemGetInfo deltas are not tracked; this reports free/total. */
int gsn_device_memory_info(size_t* free_bytes, size_t* total_bytes);

/* Per-stage kernel timing (no counterpart in the reference, which times only whole calls with
 * cv::TickMeter). While enabled, every stage records a HIP event pair on the context stream;
 * gsn_context_stage_times synchronises, returns the summed milliseconds and the number of timed
 * launches of each stage, and clears the record. Stage order:
 *   0 project, 1 scan, 2 depth sort, 3 record emit, 4 tile sort, 5 tile ranges,
 *   6 raster forward, 7 image gradients, 8 raster backward, 9 per-Gaussian gradients (+Adam),
 *   10 Adam, 11 density control (count, write, history remap, opacity cap), 12 KNN,
 *   13 statistics accumulation (per-view density and feature accumulators). */
#define GSN_STAGE_COUNT 14
int gsn_context_set_timing(gsn_context* context, int enabled);
/* Times only the stages whose bit (1 << stage) is set in stage_mask (0 disables timing). Each
 * event record costs the stream a short idle gap, so a throughput measurement times only the
 * stages it needs. No counterpart in the reference. */
int gsn_context_set_timing_mask(gsn_context* context, unsigned stage_mask);
int gsn_context_stage_times(gsn_context* context, double* milliseconds, int* launches, int stages);
/* Raster pair counting (diagnostic, no counterpart in the reference). While enabled, every raster
 * forward adds the number of (pixel, splat) pairs it evaluates (pixel not yet done, splat's band
 * kept by the conservative band mask) to a device counter of the context; the backward replays the
 * same pairs. gsn_context_raster_pairs synchronises, returns the count since the last read and
 * clears it. This is the operation count of the rasteriser's VALU roofline. */
int gsn_context_count_raster_pairs(gsn_context* context, int enabled);
int gsn_context_raster_pairs(gsn_context* context, unsigned long long* pairs);

/* Overlap for training loops: arms the next gsn_render_tiles / gsn_render_tiles_stored of the
 * context's last projection to also queue the splat backward's preparation (a packed copy of the
 * projected splats and the two work schedules, which read only the projection and tile lists) on a
 * second, default-priority stream of the context. It then runs beside the forward raster and the
 * image gradient, and the next gradient entry point of that projection uses it instead of
 * preparing again. Every gradient entry point, the next projection and gsn_context_synchronize
 * wait for the queued work first. Until one of them has run, the projection's output buffers must
 * stay allocated. One-shot: a render it does not arm does no extra work. Results are identical
 * with and without it. No counterpart in the reference (one stream throughout). */
int gsn_context_prefetch_backward(gsn_context* context);
/* Diagnostic: the number of gradient calls on this context that used work queued by an armed
 * render. No counterpart in the reference. */
int gsn_context_prefetch_reused(gsn_context* context, unsigned long long* calls);

/* Diagnostic: the wave64 reduction used by the raster backward, on one wave.
 * input float[64][12] -> output float[64][3]; lane l receives the sums over all lanes of
 * fields 3*(l/16) .. 3*(l/16)+2. No counterpart in the reference. */
int gsn_selftest_wave_sum12(gsn_context* context, const float* input, float* output);

/* Diagnostic: the rasterisers' per-tile band mask. splats float[count][6] = (mean x, mean y,
 * conic a, b, c, opacity) in the layout of the projection's means / conic_opacities; origins
 * int[count][2] = tile pixel origin (x0, y0); masks int[count]: bit b set when rows
 * y0+4b .. y0+4b+3 of the 16x16 tile may hold a pixel with alpha >= 1/255. No counterpart in the
 * reference (its rasteriser tests every pixel of the tile). */
int gsn_selftest_band_mask(gsn_context* context, const float* splats, const int* origins, int count, int* masks);

/* Diagnostic: the library's stable radix sort of (uint32 key, uint32 value) pairs by key bits
 * [0, bits) (1 <= bits <= 32), as the depth order, the tile lists and the KNN use it: keys and
 * values (device, count each) are consumed, the sorted pairs go to keys_out / values_out (device,
 * distinct from the inputs). Its counterpart in the reference is the CUB
 * DeviceRadixSort::SortPairs call of renderer.cu:546-563. */
int gsn_selftest_sort_pairs(gsn_context* context, uint32_t* keys, uint32_t* values, int count, int bits,
    uint32_t* keys_out, uint32_t* values_out);

/* Events on the context stream, for timing. */
int gsn_event_create(void** event);
int gsn_event_destroy(void* event);
int gsn_event_record(gsn_context* context, void* event);
int gsn_event_elapsed_ms(void* start, void* end, float* milliseconds);

/* ---- renderer ---------------------------------------------------------------------------- */

/* Projection + SH colour + tile rectangle + record count per Gaussian.
 * Replaces Renderer::ProjectGaussians / ProjectGaussiansKernel (renderer.cu:374-444, 108-221).
 * record_offsets[N+1] receives per-Gaussian record counts (0 when culled) and 0 at [N];
 * other outputs of culled Gaussians are left unchanged. */
int gsn_project_gaussians(gsn_context* context, const float* params, int count,
    const gsn_camera* camera, int sh_degree, float* means, float* conic_opacities,
    float* colors, float* depths, int* tile_rects, int* record_offsets);

/* In-place exclusive scan of record_offsets[N+1]; the total R is copied to *record_count
 * (waits for that copy). Replaces the DeviceScan + readback at renderer.cu:465-487.
 * When record_offsets and count are those of the last gsn_project_gaussians call on this
 * context, the call also queues the depth ordering of the visible Gaussians (using that call's
 * depths) behind the copy, so the GPU sorts while the caller reads R; gsn_build_tile_lists with
 * the same depths, record_offsets and count then reuses it. The depths must not change in
 * between. */
int gsn_scan_record_offsets(gsn_context* context, int* record_offsets, int count,
    int* record_count);

/* Tile keys, stable sort by (tile, depth bits), per-tile ranges.
 * Replaces WriteTileRecordsKernel, DeviceRadixSort::SortPairs and IdentifyTileOffsetsKernel
 * (renderer.cu:223-271, 534-569). Output order equals a stable sort of the reference keys
 * (tile << 32 | float_bits(depth)) over records enumerated in Gaussian-id order.
 * sorted_keys[R] is optional (NULL to skip). Reuses the depth order queued by
 * gsn_scan_record_offsets for the same depths / record_offsets / count, else computes it. */
int gsn_build_tile_lists(gsn_context* context, const gsn_camera* camera, int count,
    const int* tile_rects, const float* depths, const int* record_offsets, int record_count,
    int* tile_gaussian_ids, int64_t* tile_offsets, uint64_t* sorted_keys);

/* Tile rasteriser forward, front-to-back alpha blending.
 * Replaces Renderer::RenderTiles / RenderTilesKernel (renderer.cu:572-590, 273-338).
 * background points to 3 host floats. */
int gsn_render_tiles(gsn_context* context, const gsn_camera* camera,
    const int* tile_gaussian_ids, const int64_t* tile_offsets, const float* means,
    const float* conic_opacities, const float* colors, const float* background, float* image);

/* ---- solver ------------------------------------------------------------------------------ */

/* dL/dimage for L = 0.8*mean|I-T| + 0.2*(1-mean SSIM), 11-tap sigma 1.5 window, zero padding.
 * Replaces Solver::ComputeImageGradients / ComputeImageGradientsKernel (solver.cu:735-759, 26-168). */
int gsn_image_gradients(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, float* image_gradients);

/* Loss value and image metrics (no counterpart in the reference, which never evaluates the
 * loss on the training path, solver.cu:162-166). out[0] = 0.8*L1 + 0.2*(1-SSIM), out[1] = mean
 * |I-T|, out[2] = mean SSIM, out[3] = mean squared error of the image clamped to [0,1] (for
 * PSNR). Synchronises the stream. */
int gsn_image_metrics(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, double* out);

/* Held-out evaluation of a render (replaces the scoring of reference demo/render.cpp:130-140):
 * the image is clamped to [0, 1], scaled by 255 in FP32 and rounded half to even to 8 bits
 * (written to `quantized`, HWC, when not null), and compared with the 8-bit target.
 * out[0] = sum of squared differences over all pixels and channels (exact),
 * out[1] = PSNR in dB, -10 log10(out[0] / (3 W H 255^2)), +inf when out[0] = 0. Synchronises. */
int gsn_image_quantized_error(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, uint8_t* quantized, double* out);

/* Allocation-free gsn_image_metrics: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, evaluation loops that must not call
 * hipMalloc). Same kernels and results (bit for bit) as gsn_image_metrics, but the per-block
 * partial sums live in a file-scope device array of 3 MiB per device (131,072 blocks of 32x16
 * pixels per channel: images up to about 22 million pixels; larger ones return
 * GSN_ERROR_CAPACITY), and the four results are written to device_out (4 doubles of device
 * memory) without synchronising the stream. Calls that overlap on one device share that array
 * and race; calls on one context's stream are ordered and safe. */
int gsn_image_metrics_static(gsn_context* context, int width, int height, const float* image,
    const uint8_t* target, double* device_out);

/* Splat backward (recomputes transmittance front to back from the final colour), projection
 * statistics and the screen-to-parameter chain rule.
 * Replaces ComputeSplatGradientsKernel, CollectProjectionStatisticsKernel and
 * MapGaussianGradientsKernel<D> (solver.cu:837-899, 325-623).
 * gradients[58][N]: parameter gradients (rows of unobserved Gaussians are zero).
 * screen_gradients[11][N] (optional): dmean2D(2), dconic(3), dlogit, dcolor(3), |dmean2D|(2).
 * statistics[2][N]: (|dmean2D| scaled by W/2,H/2; observed flag).
 * absolute_statistics[2][N]: row 0 = norm of the per-pixel absolute dmean2D sums. */
int gsn_gaussian_gradients(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, const float* params, const int* tile_gaussian_ids,
    const int64_t* tile_offsets, int record_count, const int* record_offsets,
    const int* tile_rects, const float* means, const float* conic_opacities,
    const float* colors, const float* image, const float* image_gradients, float* gradients,
    float* screen_gradients, float* statistics, float* absolute_statistics);

/* Adam moments. The reference keeps two [58][N] arrays (Solver's private first_moments_ /
 * second_moments_, solver.h:60-63); here they are one blocked buffer, the optimiser's own state:
 * block b holds Gaussians [64b, 64b+64), and inside it each of the 58 rows holds 64 (first,
 * second) float pairs, so the pair of Gaussian i in row r is at float
 * ((i / 64) * 58 + r) * 128 + (i % 64) * 2. gsn_moment_floats(N) is its size in floats (N rounded
 * up to 64). gsn_moments_unpack / gsn_moments_pack convert to and from two [58][N] arrays (no
 * counterpart in the reference: tests and checkpoints). */
size_t gsn_moment_floats(int count);
int gsn_moments_unpack(gsn_context* context, int count, const float* moments, float* first, float* second);
int gsn_moments_pack(gsn_context* context, int count, const float* first, const float* second, float* moments);

/* Same as gsn_gaussian_gradients followed by gsn_adam_update, fused so that parameter
 * gradients never go to memory. Replaces Solver::Iterate's gradient + update stages
 * (solver.cu:702-720, 837-929). moments: blocked, gsn_moment_floats(count) floats. */
int gsn_gaussian_gradients_adam(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets,
    int record_count, const int* record_offsets, const int* tile_rects, const float* means,
    const float* conic_opacities, const float* colors, const float* image,
    const float* image_gradients, float* moments, float position_learning_rate,
    float first_correction, float second_correction, float* statistics, float* absolute_statistics);

/* gsn_gaussian_gradients_adam followed by gsn_project_gaussians of the updated parameters for
 * next_camera (same SH degree), fused: the pass projects each Gaussian from the values it has
 * just written, so the next iteration's projection reads no parameter rows. The next_* buffers
 * receive exactly what gsn_project_gaussians would write (record counts in next_record_offsets,
 * ready for gsn_scan_record_offsets) and become the context's last projection.
 * next_record_offsets must not be record_offsets; the other next_* buffers may be the current
 * projection's (the gradient pass does not read them). Replaces Solver::Iterate
 * (solver.cu:702-720) and the projection of the following Renderer::Render (renderer.cu:374-444)
 * when nothing changes the parameters in between (density control and opacity caps do).
 * gradient_sums, absolute_gradient_sums, observation_counts: all null, or the density
 * accumulators, to which the pass adds this view's statistics exactly as a following
 * gsn_density_accumulate would (the statistics are written either way); a caller that passes
 * them does not call gsn_density_accumulate for this view (replaces
 * AccumulateRoundStatisticsKernel, density_controller.cu:280-295, for this view). */
int gsn_gaussian_gradients_adam_project(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets,
    int record_count, const int* record_offsets, const int* tile_rects, const float* means,
    const float* conic_opacities, const float* colors, const float* image,
    const float* image_gradients, float* moments, float position_learning_rate,
    float first_correction, float second_correction, float* statistics, float* absolute_statistics,
    const gsn_camera* next_camera, float* next_means, float* next_conic_opacities, float* next_colors,
    float* next_depths, int* next_tile_rects, int* next_record_offsets, float* gradient_sums,
    float* absolute_gradient_sums, uint32_t* observation_counts);

/* Dense Adam over all 58*N entries with the reference's per-group learning rates.
 * Replaces UpdateGaussiansKernel (solver.cu:651-689, 901-929). */
int gsn_adam_update(gsn_context* context, int count, const float* gradients, float* params,
    float* moments, float position_learning_rate, float first_correction, float second_correction);

/* ---- stored-state rasteriser variant (no counterpart in the reference) -------------------
 * For the peak-memory comparison with the recompute form above, which stores nothing per pixel.
 * The forward additionally writes, per pixel, the final transmittance (float[H][W]) and the
 * tile-relative index + 1 of the last splat blended (int[H][W], 0 if none): 8 bytes per pixel.
 * The backward then walks each tile's records back to front from the last contributor,
 * recovering T_k = T_{k+1} / (1 - alpha_k) and the colour behind each splat from T_N * background
 * (the form of the original 3DGS rasteriser). Same outputs and partial layout as the recompute
 * entries; background points to 3 host floats and must equal the forward's. */
int gsn_render_tiles_stored(gsn_context* context, const gsn_camera* camera,
    const int* tile_gaussian_ids, const int64_t* tile_offsets, const float* means,
    const float* conic_opacities, const float* colors, const float* background, float* image,
    float* final_transmittance, int32_t* last_contributor);
int gsn_gaussian_gradients_stored(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, const float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets,
    int record_count, const int* record_offsets, const int* tile_rects, const float* means,
    const float* conic_opacities, const float* colors, const float* final_transmittance,
    const int32_t* last_contributor, const float* background, const float* image_gradients,
    float* gradients, float* screen_gradients, float* statistics, float* absolute_statistics);
int gsn_gaussian_gradients_adam_stored(gsn_context* context, const gsn_camera* camera, int count,
    int sh_degree, float* params, const int* tile_gaussian_ids, const int64_t* tile_offsets,
    int record_count, const int* record_offsets, const int* tile_rects, const float* means,
    const float* conic_opacities, const float* colors, const float* final_transmittance,
    const int32_t* last_contributor, const float* background, const float* image_gradients,
    float* moments, float position_learning_rate, float first_correction, float second_correction,
    float* statistics, float* absolute_statistics);

/* Gathers Adam moments (blocked buffers, distinct) by parent id (parent < 0 or >= previous_count
 * -> zero). parent_ids may be NULL (all zero). Replaces RemapGaussianHistoryKernel
 * (solver.cu:625-649). */
int gsn_remap_history(gsn_context* context, const int* parent_ids, int previous_count,
    int count, const float* previous_moments, float* moments);

/* Zeroes both moments of rows [row, row+rows) of all `count` Gaussians (opacity history reset,
 * solver.cu:722-733). */
int gsn_zero_rows(gsn_context* context, float* moments, int count, int row, int rows);

/* ---- density control ------------------------------------------------------------------- */

/* Per-Gaussian decision {keep, clone, split, split-kept} and its exclusive prefix.
 * Replaces CountDensityChangesKernel + DeviceScan::ExclusiveScan<int4> + the totals readback
 * (density_controller.cu:60-116, 482-541). decisions/offsets are int4[N+1]; totals (host, 4
 * ints) = offsets[N]. Synchronises the stream. */
int gsn_density_count(gsn_context* context, const float* params, int count,
    const float* gradient_sums, const float* absolute_gradient_sums,
    const uint32_t* observation_counts, float gradient_threshold, float split_threshold,
    float min_opacity, float max_world_size, int prune_large,
    float absolute_gradient_threshold, int32_t* decisions, int32_t* offsets, int32_t* totals);

/* Writes the compacted parameter buffer [keep | clone | split child 0 | split child 1].
 * Split children move by R*(s (.) n), n ~ N(0, I) from Philox4x32-10 keyed by (seed, counter),
 * and take log-scale log(s/1.6). Also writes history parent ids (kept -> old id, else -1),
 * parent ids (old id for every output) and child flags (1 for clone/split outputs).
 * Replaces GenerateDensitySamplesKernel, WriteDensityGaussiansKernel and
 * FinishDensityControlKernel (density_controller.cu:138-278). Any of the three id/flag
 * outputs may be NULL. */
int gsn_density_write(gsn_context* context, const float* params, int count,
    const int32_t* decisions, const int32_t* offsets, const int32_t* totals, uint64_t seed,
    uint64_t counter, float* out_params, int out_count, int32_t* history_parent_ids,
    int32_t* parent_ids, float* child_flags);

/* The split-jitter generator on its own: out[i] = i-th standard normal for (seed, counter). */
int gsn_density_normals(gsn_context* context, uint64_t seed, uint64_t counter, int64_t n,
    float* out);

/* sums += statistics[0], absolute_sums += absolute_statistics[0], counts += statistics[1]
 * for ids < count. Replaces AccumulateRoundStatisticsKernel (density_controller.cu:280-295). */
int gsn_density_accumulate(gsn_context* context, const float* statistics,
    const float* absolute_statistics, int count, float* gradient_sums,
    float* absolute_gradient_sums, uint32_t* observation_counts);

/* Per-observation accumulation of |(gx, gy)| for radius > 0. gaussian_ids may be NULL
 * (identity). Replaces AccumulateDensityStatisticsKernel (density_controller.cu:39-58). */
int gsn_density_accumulate_observations(gsn_context* context, const float* gradients,
    const float* radii, const int32_t* gaussian_ids, int observations, int count,
    float* gradient_sums, uint32_t* observation_counts);

/* logit <- min(logit, max_opacity_logit) on row 9 of a [58][stride] buffer, first count ids.
 * Replaces ResetDensityOpacityKernel (density_controller.cu:297-308). */
int gsn_density_reset_opacity(gsn_context* context, float* params, int stride, int count,
    float max_opacity_logit);

/* Zeroes the accumulators for the first count Gaussians. */
int gsn_density_clear(gsn_context* context, float* gradient_sums,
    float* absolute_gradient_sums, uint32_t* observation_counts, int count);

/* ---- learned densification inputs and policy (no counterpart in the reference) ----------- */

/* gsn_density_count with the gradient-threshold selection replaced by a proposal: for each
 * Gaussian the largest of logits[N][3] (keep, clone, split; ties and NaN favour keep, then clone)
 * chooses the densification; pruning (opacity, oversize) stays the reference rule
 * (density_controller.cu:95-115). Same outputs and synchronisation as gsn_density_count. */
int gsn_density_count_proposed(gsn_context* context, const float* params, int count,
    const float* gradient_sums, const float* absolute_gradient_sums,
    const uint32_t* observation_counts, const float* proposal_logits, float split_threshold,
    float min_opacity, float max_world_size, int prune_large, int32_t* decisions,
    int32_t* offsets, int32_t* totals);

/* Zeroes a [6][stride] feature state (max gradient, max absolute gradient, depth sum, depth min,
 * depth max, max footprint radius). */
int gsn_density_features_clear(gsn_context* context, float* state, int stride);

/* Folds one view into the feature state of every observed Gaussian (statistics[1][id] != 0):
 * statistics / absolute_statistics are the [2][N] outputs of the gradient pass, depths and
 * conic_opacities the projection's. */
int gsn_density_features_accumulate(gsn_context* context, const float* statistics,
    const float* absolute_statistics, const float* depths, const float* conic_opacities, int count,
    float* state, int stride);

/* features[N][16] for the proposal network: gradient means / threshold (from the density
 * accumulators, as the rule uses them), log1p(max gradients / threshold), log1p(observations),
 * descending log-scales - log(split_threshold), opacity, log(depth mean/min/max / extent),
 * SH DC, log1p(max footprint radius); observation-derived entries are 0 when unobserved. */
int gsn_density_features(gsn_context* context, const float* params, int count,
    const float* gradient_sums, const float* absolute_gradient_sums,
    const uint32_t* observation_counts, const float* state, int stride,
    float gradient_threshold, float split_threshold, float scene_extent, float* features);

/* Training labels from the outcome of the previous density step. decisions/offsets/totals are
 * that step's gsn_density_count outputs over previous_count Gaussians; params and
 * observation_counts describe the count Gaussians it produced, after the interval that followed.
 * labels[previous_count]: -1 pruned, 0 kept, 1 cloned with both copies useful, 2 split with both
 * children useful (useful: observed in the interval and opacity >= min_opacity); a densified
 * Gaussian whose outputs were not all useful is labelled 0. */
int gsn_density_outcomes(gsn_context* context, const int32_t* decisions, const int32_t* offsets,
    const int32_t* totals, int previous_count, const float* params,
    const uint32_t* observation_counts, int count, float min_opacity, int32_t* labels);

/* ---- initialisation ---------------------------------------------------------------------- */

/* Three nearest neighbours of every point (ties broken by smaller index) through a linear BVH
 * over Morton-ordered points. neighbors[n][3] (optional) receives indices (-1 when fewer than
 * 3 others exist); mean_dist2[n] receives the mean squared distance (reference rule for n < 4).
 * points is device float[n][3]. Replaces ComputeMeanDist2Cuda's device part (knn.cu:180-255). */
int gsn_knn3(gsn_context* context, const float* points, int n, int32_t* neighbors,
    float* mean_dist2);

/* ---- learned densification proposal (no counterpart in the reference) ------------------ */

/* Number of FP32 master parameters (5443: W1[64][16], b1[64], W2[64][64], b2[64], W3[3][64],
 * b3[3]) and of BF16 weight entries (7168: W1[64][16], W2[64][64], W3 padded to [32][64]) of
 * the 16 -> 64 -> 64 -> 3 proposal network. */
int gsn_mlp_sizes(int* parameters, int* bf16_weights);

/* Rounds the FP32 master weights to the BF16 copy the MFMA kernels read (round to nearest
 * even), zeroing the W3 padding rows. weights is uint16_t[7168] holding BF16 bit patterns. */
int gsn_mlp_sync_weights(gsn_context* context, const float* params, uint16_t* weights);

/* logits[n][3] = W3 relu(W2 relu(W1 x + b1) + b2) + b3 for features[n][16] (FP32 in HBM,
 * rounded to BF16 for the MFMA operands; hidden activations are rounded to BF16 after the
 * FP32 bias + ReLU; accumulation is FP32). Biases come from the FP32 parameter block. */
int gsn_mlp_forward(gsn_context* context, const float* features, int count,
    const uint16_t* weights, const float* params, float* logits);

/* One full-batch training step on the weighted mean softmax cross-entropy over labels
 * (0 keep, 1 clone, 2 split): gradients of all 5443 parameters (MFMA forward and backward,
 * per-workgroup sums reduced in a fixed order), then Adam (beta 0.9 / 0.999, eps 1e-8, bias
 * corrections for `step`) on params/first/second when learning_rate > 0, and a refresh of the
 * BF16 copy. sample_weights may be NULL (all 1). gradients_out[5443] and loss_out (host; the
 * loss before the update) are optional; loss_out synchronises the stream. */
int gsn_mlp_train_step(gsn_context* context, const float* features, const int32_t* labels,
    const float* sample_weights, int count, float* params, float* first, float* second,
    uint16_t* weights, float learning_rate, int step, float* gradients_out, double* loss_out);

#ifdef __cplusplus
}
#endif

#endif /* GSN_HIP_H */
