/* fastwarp_hip.h — C ABI of the MI355X (gfx950) registration engine.
 *
 * This is the drop-in boundary of fastwarp's GPU registration hot path. The reference drives its
 * device through three Rust traits:
 *   - fastwarp_reg::Backend          crates/fastwarp-reg/src/backend.rs:39-305 (the SyN loop's ops)
 *   - fastwarp_reg::LinearEngine     crates/fastwarp-reg/src/linear.rs:288-320
 *   - fastwarp_reg::LinearMetric     crates/fastwarp-reg/src/optimizer.rs:60-65
 * and implements them with wgpu in crates/fastwarp-gpu/src/wgpu_backend.rs (`WgpuBackend`) and
 * crates/fastwarp-gpu/src/wgpu_backend/mi.rs (`GpuMiLevel`). Every entry point below replaces one
 * trait method (cited per function), with opaque handles for the trait's associated
 * `Image`/`Field` types. Plain pointers and sizes only; no C++ or torch types cross this line.
 *
 * Conventions
 *   - One fw_ctx is one HIP device, one stream and one buffer pool. Calls on one context are
 *     serialised by the caller (the trait's ops are issued from one thread, backend.rs:36-37);
 *     different contexts may be used from different threads.
 *   - Every function returning int returns FW_OK (0) or a negative status; fw_last_error()
 *     gives the message ("what: message", as RegError::Invalid prints).
 *   - Images are f32 on the device, x-fastest (wgpu_backend.rs:120-135). Fields are three f32
 *     planes (x, y, z) of LPS millimetre displacements (wgpu_backend.rs:159-170). Host-side
 *     arrays crossing the ABI are f64: images as n values, fields as 3n interleaved (dx,dy,dz),
 *     as Volume<f64> and VectorField are on the host side of the reference.
 *   - Ownership follows the trait: functions documented as consuming a handle take it over
 *     (it must not be used or freed afterwards), exactly where the trait takes `Self::Field` by
 *     value (smooth_field, invert, invert_pair, warp_mask_through_field's `product`).
 */
#ifndef FASTWARP_HIP_H
#define FASTWARP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FW_OK 0
#define FW_ERR_INVALID (-1)   /* invalid argument or state (RegError::Invalid) */
#define FW_ERR_DEVICE (-2)    /* HIP runtime error or device allocation failure */
#define FW_ERR_CANCELLED (-3) /* the cancel flag was set (RegError::Cancelled) */

typedef struct fw_ctx fw_ctx;
typedef struct fw_image fw_image;
typedef struct fw_field fw_field;
typedef struct fw_mi_level fw_mi_level;

/* An ITK image grid in LPS mm: p = origin + direction * diag(spacing) * index; direction holds
 * the axis directions as columns, row-major direction[3*row + axis] (grid.rs:26-44). */
typedef struct {
  int64_t size[3];
  double spacing[3];
  double origin[3];
  double direction[9];
} fw_grid;

/* y = M (x - c) + c + t (transform/affine.rs:34-45), matrix row-major. */
typedef struct {
  double matrix[9];
  double translation[3];
  double center[3];
} fw_affine;

/* InvertParams (field_ops.rs:276-295); SyN uses 20, 0.001, 0.1, 1. */
typedef struct {
  uint32_t max_iterations;
  double mean_error_tolerance;
  double max_error_tolerance;
  int32_t enforce_boundary_condition;
} fw_invert_params;

/* InvertStats (field_ops.rs:298-307). */
typedef struct {
  uint32_t iterations;
  double mean_error_norm;
  double max_error_norm;
} fw_invert_stats;

/* MiBinning (metric/mi.rs:121-143) of one linear-stage level. */
typedef struct {
  int64_t bins;
  double fixed_min, fixed_max, moving_min, moving_max;
  double fixed_bin_size, moving_bin_size, fixed_normalized_min, moving_normalized_min;
} fw_mi_binning;

/* LinearMovingTransform (transform/param.rs:306-315): x -> outer(optimised(x)). kind 0 is
 * Euler3D (6 parameters, matrix Rz Rx Ry), kind 1 is affine (12: matrix row-major, translation). */
typedef struct {
  fw_affine outer;
  int32_t kind;
  double parameters[12];
  double center[3];
} fw_linear_transform;

/* ---- context (WgpuBackend::new, release_idle_buffers backend.rs:50) ---- */
int fw_ctx_create(int device, fw_ctx** out);
/* Frees the context's device memory. Every image, field and MI level made on it must be freed
 * before (their buffers return to its pool); a handle freed after this call is undefined. */
void fw_ctx_destroy(fw_ctx* ctx);
const char* fw_last_error(const fw_ctx* ctx);
/* Backend::release_idle_buffers (backend.rs:48-50). */
int fw_release_idle(fw_ctx* ctx);
/* Sets one launch-shape choice of the context (no counterpart in the reference, whose shaders
 * have fixed workgroup shapes): key "<knob>" or "<knob>@<class>", class = floor(log2(size)) of the
 * grid's voxels or the level's samples; a negative value removes the entry. Knobs: cc.tile_z
 * (1..256), cc.min_blocks, inv.max_blocks, mi.threads (256, 512, 1024), mi.max_blocks,
 * mi.batch_blocks (the cap of the passes that bin several transforms; mi.max_blocks's by default),
 * mi.deriv_blocks, mi.fused_slots (passes of at most this many blocks add their slots and finish in
 * one launch; 0 never), gauss.one_pass (0, 1), deriche.lds (0, 1: x lines of the level smoothing
 * staged in LDS; default 1), batch.graphs (0, 1: the batched SyN replays each
 * iteration's update launches from a HIP graph; default 0). A context starts from the table the
 * project's autotuner wrote (fastwarp_amd/tuning/gfx950.json, or the file FASTWARP_TUNING names);
 * results change at most by rounding between choices (batch.graphs not at all). */
int fw_set_tuning(fw_ctx* ctx, const char* key, int64_t value);
/* Device bytes the context's pool holds (in use plus idle) and its peak since creation. */
int fw_memory_stats(fw_ctx* ctx, uint64_t* held, uint64_t* peak);

/* ---- images (Backend::upload_image backend.rs:53-57) ---- */
int fw_image_upload(fw_ctx* ctx, const fw_grid* grid, const double* data, fw_image** out);
int fw_image_upload_f32(fw_ctx* ctx, const fw_grid* grid, const float* data, fw_image** out);
int fw_image_download(fw_ctx* ctx, const fw_image* image, double* out);
/* The pipeline's isotropic resample (crates/fastwarp-reg/src/register.rs:112-142
 * `resample_isotropic`, which `register` applies to every raw input when given an iso_mm): a new
 * image on the grid of spacing mm with the same origin and direction and
 * floor((n - 1) * spacing / mm + 1e-9) + 1 voxels per axis, trilinear from the source indices
 * min(i * mm / spacing, n - 1); a copy when every axis already has spacing mm (within 1e-6). On
 * the device, in f32 with the index splits computed in f64. A caller that registers at iso_mm
 * resamples each raw image with it before fw_register_resident. */
int fw_resample_isotropic(fw_ctx* ctx, const fw_image* image, double mm, fw_image** out);
int fw_image_grid(const fw_image* image, fw_grid* out);
void fw_image_free(fw_ctx* ctx, fw_image* image);

/* ---- fields (Backend::zero_field/upload_field/download_field/field_grid/copy_field
 *      backend.rs:59-72) ---- */
int fw_field_zero(fw_ctx* ctx, const fw_grid* grid, fw_field** out);
int fw_field_upload(fw_ctx* ctx, const fw_grid* grid, const double* aos, fw_field** out);
int fw_field_download(fw_ctx* ctx, const fw_field* field, double* aos);
int fw_field_grid(const fw_field* field, fw_grid* out);
int fw_field_copy(fw_ctx* ctx, const fw_field* field, fw_field** out);
void fw_field_free(fw_ctx* ctx, fw_field* field);

/* Backend::warp_through_field (backend.rs:80-86): out(x) = src(affine(q + pre(q))), q = x + field(x);
 * pre and affine may be NULL. */
int fw_warp(fw_ctx* ctx, const fw_image* src, const fw_field* field, const fw_field* pre, const fw_affine* affine,
            fw_image** out);
/* Backend::warp_mask_through_field (backend.rs:95-102): nearest neighbour, {0,1}; `product`
 * (consumed, may be NULL) is multiplied in and its buffer returned. */
int fw_warp_mask(fw_ctx* ctx, const fw_image* mask, const fw_field* field, const fw_field* pre, const fw_affine* affine,
                 fw_image* product, fw_image** out);

/* Backend::cc_value_derivative_pair(_masked)(_with_norms) (backend.rs:121-177): both directions of
 * ANTs neighbourhood CC from one set of window sums. `valid` may be NULL (no masks). values[0] is
 * (fixed, moving), values[1] (moving, fixed); d_fixed / d_moving are new fields. norms (may be
 * NULL) receives the l2 norms of both derivatives. */
int fw_cc_pair(fw_ctx* ctx, const fw_image* fixed, const fw_image* moving, const fw_image* valid, uint32_t radius,
               double values[2], double norms[2], fw_field** d_fixed, fw_field** d_moving);

/* No counterpart in the reference: fw_cc_pair (unmasked) for callers that cannot allocate (code
 * under HIP graph capture, serving loops that forbid hipMalloc). Every buffer is the caller's
 * device memory: fixed and moving (n f32 each, x-fastest, on `grid`), d_fixed and d_moving (3n f32,
 * three planes each, written), values (2 f64, written: values[0] = values[1], as fw_cc_pair's).
 * The work is enqueued on `stream` (a hipStream_t, NULL for the null stream) and the call returns
 * without waiting; nothing is allocated. Block partials live in a static per-device scratch of
 * 137,224 bytes, which limits the grid to 16,384 CC blocks of 64 x 8 x (at most 16) voxels (the
 * 1 mm grid needs 1008); a larger grid is rejected with FW_ERR_INVALID. Calls that overlap on one
 * GPU race on that scratch. Results are the same bits as fw_cc_pair's on the same images. Errors:
 * the status, and fw_last_error(NULL) for the text. */
int fw_cc_pair_noalloc(const fw_grid* grid, const float* fixed, const float* moving, uint32_t radius, float* d_fixed,
                       float* d_moving, double* values, void* stream);

/* Backend::mi_value_derivative_pair(_masked)(_with_norms) (backend.rs:185-243): dense Mattes MI in
 * both directions; masks may be NULL. */
int fw_mi_dense_pair(fw_ctx* ctx, const fw_image* fixed, const fw_image* moving, const fw_image* fixed_mask,
                     const fw_image* moving_mask, uint32_t bins, double values[2], double norms[2], fw_field** d_fixed,
                     fw_field** d_moving);

/* Backend::smooth_field (backend.rs:248): consumes `field`; variance <= 0 returns it unchanged. */
int fw_smooth_field(fw_ctx* ctx, fw_field* field, double variance, fw_field** out);
/* Backend::max_scaled_norm_pair (backend.rs:251-260); b may be NULL (out[1] untouched). */
int fw_max_scaled_norm_pair(fw_ctx* ctx, const fw_field* a, const fw_field* b, double out[2]);
/* Backend::l2_norm (backend.rs:263). */
int fw_l2_norm(fw_ctx* ctx, const fw_field* field, double* out);
/* Backend::scale (backend.rs:266). */
int fw_scale(fw_ctx* ctx, fw_field* field, double s);
/* Backend::axpy (backend.rs:269). */
int fw_axpy(fw_ctx* ctx, fw_field* acc, double a, const fw_field* x);
/* Backend::compose (backend.rs:273-277): out(x) = warping(x) + displacement(x + warping(x)). */
int fw_compose(fw_ctx* ctx, const fw_field* displacement, const fw_field* warping, fw_field** out);
/* Backend::invert (backend.rs:281-286): consumes `initial`. */
int fw_invert(fw_ctx* ctx, const fw_field* field, fw_field* initial, const fw_invert_params* params, fw_field** out,
              fw_invert_stats* stats);
/* Backend::invert_pair (backend.rs:290-301): consumes both; new_inverse = invert(field, inverse),
 * then new_field = invert(new_inverse, field). */
int fw_invert_pair(fw_ctx* ctx, fw_field* field, fw_field* inverse, const fw_invert_params* params,
                   fw_field** new_inverse, fw_invert_stats* inverse_stats, fw_field** new_field,
                   fw_invert_stats* field_stats);
/* Backend::resample_field (backend.rs:304). */
int fw_resample_field(fw_ctx* ctx, const fw_field* field, const fw_grid* grid, fw_field** out);

/* ---- linear-stage Mattes MI (LinearEngine::prepare_level linear.rs:298-319, LinearMetric
 *      optimizer.rs:60-65, GpuMiLevel wgpu_backend/mi.rs:144-524) ---- */
/* points: 3m physical LPS coordinates, fixed_bin: m fixed-axis bins (PreparedSamples,
 * linear.rs:240-279). The moving image handle must outlive the level. */
int fw_mi_level_create(fw_ctx* ctx, const fw_image* moving, const double* points, const uint32_t* fixed_bin, int64_t m,
                       const fw_mi_binning* binning, fw_mi_level** out);
int fw_mi_level_value(fw_mi_level* level, const fw_linear_transform* t, double* value);
/* GpuMiLevel::histogram (crates/fastwarp-gpu/src/wgpu_backend/mi.rs:312-377): the merged histogram
 * of one transform before the finish: joint[bins * bins] (fixed bin major; the B-spline weights
 * summed, in samples), fixed_marginal[bins] (sample counts) and the valid sample count. */
int fw_mi_level_histogram(fw_mi_level* level, const fw_linear_transform* t, double* joint, double* fixed_marginal,
                          int64_t* valid);
/* values[i] = the value of transforms[i], i < count, computed in as few histogram passes as the
 * device allows (up to four transforms share one pass over the samples); each equals
 * fw_mi_level_value of that transform. No counterpart in the reference: the line search's
 * speculative probes (two golden-section steps per pass) use it. */
int fw_mi_level_values(fw_mi_level* level, const fw_linear_transform* transforms, int64_t count, double* values);
/* derivative receives 6 or 12 values. */
int fw_mi_level_value_and_derivative(fw_mi_level* level, const fw_linear_transform* t, double* value,
                                     double* derivative);
void fw_mi_level_free(fw_mi_level* level);

/* ---- level images on the device (pyramid.rs:429-465 smooth_recursive_gaussian; the reference
 *      smooths on the host in f64 and uploads, this engine smooths the uploaded image in place of
 *      that: same filter, axes z, x, y, f32 arithmetic) ---- */
/* sigma_mm per axis, each > 0 (or all three 0 for a copy); every axis needs 4 voxels. */
int fw_smooth_image(fw_ctx* ctx, const fw_image* src, const double sigma_mm[3], fw_image** out);
/* Values of `img` at m physical points (3m f64): the nearest voxel, rounding half up, and the
 * linear interpolation (mi.rs:161-207, linear.rs:251-279); NaN outside. Either output may be NULL. */
int fw_sample_values(fw_ctx* ctx, const fw_image* img, const double* points, int64_t m, double* nearest, double* linear);
/* The winsorised copy of a raw image on the device: the 256-bin quantile window [lower, upper]
 * mapped onto [0, 1] (preprocess.rs:159-232; the histogram bins equal the f64 ones for the f32
 * intensities). Fails, as fw_register does, when a voxel is not finite. */
int fw_preprocess_image(fw_ctx* ctx, const fw_image* raw, float lower, float upper, fw_image** out);
/* The intensity-weighted centre of an image on the device, physical mm (init.rs:47-86). */
int fw_centre_of_gravity(fw_ctx* ctx, const fw_image* img, double centre[3]);

/* ---- whole registrations (stages.rs:286 register_stages_async with the HIP backend) ---- */

/* The stage plan of one full normalisation: centre of mass, Rigid, Affine (Mattes MI), SyN (CC).
 * Per-level vectors hold `levels` entries. */
typedef struct {
  int32_t linear_levels;
  uint32_t linear_iterations[8];
  uint32_t linear_shrink[8];
  double linear_sigmas[8];
  double linear_gradient_step;
  int32_t mi_bins;
  double sampling_percentage;
  uint32_t seed;
  double linear_convergence_threshold;
  int32_t linear_convergence_window;

  int32_t syn_levels;
  uint32_t syn_iterations[8];
  uint32_t syn_shrink[8];
  double syn_sigmas[8];
  double syn_gradient_step;
  double update_variance;
  double total_variance;
  int32_t cc_radius;
  int32_t channels;
  double channel_weights[8];
  double syn_convergence_threshold;
  int32_t syn_convergence_window;

  double winsorize_lower, winsorize_upper;
  int32_t run_linear; /* 0: no linear stage, SyN starts from the initial transform */
  int32_t run_syn;    /* 0: linear stages only */
  /* Where the chain starts (the reference's StagePlan::initial, crates/fastwarp-reg/src/stages.rs:74-86,
   * ANTs' --initial-moving-transform): 0 the translation between the centres of intensity mass of
   * the raw first channel, 1 the identity centred at the origin, 2 initial_affine (fixed points to
   * moving points; finite and invertible, the first linear stage is centred at its centre). */
  int32_t initial_kind;
  fw_affine initial_affine;
  /* The SyN metric of each channel (recipe.rs SynRecipe::channel_metrics): 0 CC (radius
   * cc_radius), 1 Mattes MI (mi_bins bins, dense); all 0 by default. */
  int32_t syn_channel_metrics[8];
} fw_plan;

/* Summary of one registration: the collapsed affine, per-phase seconds and per-level counts. */
typedef struct {
  fw_affine affine;
  double seconds_total, seconds_preprocess, seconds_linear, seconds_syn;
  int32_t syn_levels;
  int32_t syn_iterations[8];
  double syn_last_energy[8];
  int32_t linear_iterations[2][8];
  double linear_final_value[2];
  double rigid_parameters[6];
  double affine_parameters[12];
  /* per level: wall seconds of the level and of its iterations alone (the rest is set-up:
   * smoothing, sampling, uploads); linear index 0 is Rigid, 1 Affine */
  double syn_level_seconds[8];
  double syn_iteration_seconds[8];
  double linear_level_seconds[2][8];
  double linear_iteration_seconds[2][8];
  int64_t linear_evaluations[2][8]; /* metric evaluations, line searches included */
  int64_t syn_inversion_steps[8];    /* executed fixed-point steps of the level's four inversions per iteration, summed */
} fw_result;

/* One iteration of the last registration run on the calling thread: a CG iteration of a linear
 * level (stage 0 rigid, 1 affine; value = the metric value of the iteration, as the reference's
 * CgTrace::values, optimizer.rs:324-430), or a SyN iteration (stage 2; value = the energy, with
 * the convergence value, the two update scales and the four inversions' step counts and errors of
 * syn.rs:96-128 `SynIteration`). */
typedef struct {
  int32_t stage, level, iteration, pad;
  double value;
  double convergence_value;
  double update_scale[2];
  int32_t inversion_steps[4];
  double inversion_mean_error[4];
  double inversion_max_error[4];
} fw_trace_step;
/* Copies up to `capacity` records of the calling thread's last fw_register / fw_register_resident
 * (in stage, level, iteration order) into `out` (may be NULL with capacity 0) and returns how
 * many there are. The reference returns these traces inside its stage results
 * (crates/fastwarp-reg/src/syn.rs:96-128, optimizer.rs:324-430). */
int32_t fw_last_trace(fw_trace_step* out, int32_t capacity);

/* ---- profiling (the reference's WgpuBackend::set_profiling / take_profile,
 *      crates/fastwarp-gpu/src/profiler.rs:95-170) ---- */
/* While on, every kernel launch carries HIP events on the context's stream
 * (hipExtLaunchKernelGGL: the span is the kernel's own execution). */
int fw_set_profiling(fw_ctx* ctx, int on);
/* Times only every `every`-th launch of the kernel family `family` (NULL or "": every launch of
 * every family); the other launches are counted with their bytes but carry no events. No
 * counterpart in the reference: it keeps the cost of the events off most launches of a timed run. */
int fw_set_profiling_focus(fw_ctx* ctx, const char* family, int every);
/* Writes one line per kernel family, "name launches timed_ms algorithmic_bytes timed_launches",
 * into `text` (NUL-terminated, at most `capacity` bytes) and clears the totals; timed_ms covers
 * the timed launches only. Host spans of the drivers are listed as "host:<name>". */
int fw_take_profile(fw_ctx* ctx, char* text, size_t capacity);

/* ---- admission: device memory of one registration (the reference's capacity model,
 *      crates/fastwarp-gpu/src/capacity.rs:162-180, 511-715, and fits_gpu /
 *      registrations_at_once, crates/fastwarp-py/src/lib.rs:322-380, for this engine's buffers) ---- */
typedef struct {
  uint64_t fixed_voxels, moving_voxels;
  uint64_t smoothing_cache_bytes; /* both images of every channel at every distinct sigma */
  uint64_t linear_bytes;          /* peak of the linear chain (0 when it does not run) */
  uint64_t syn_bytes;             /* peak of the SyN stage (0 when it does not run) */
  uint64_t peak_bytes;            /* what one registration needs: the larger peak plus overhead */
  uint64_t largest_buffer_bytes;  /* the largest single allocation */
} fw_capacity;
/* Pure, no device: FW_ERR_INVALID for a bad plan or grid. */
int fw_plan_capacity(const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid, fw_capacity* out);
/* Free device memory of the context's device less 1 GiB kept in reserve, or the environment
 * override FASTWARP_GPU_MEMORY_BUDGET_MB (capacity.rs:162-180). */
int fw_memory_budget(fw_ctx* ctx, uint64_t* bytes);
/* Whether one registration of `plan` fits the budget now; how many fit at once (at most `max`). */
int fw_fits(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid, int32_t* fits);
int fw_registrations_at_once(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const fw_grid* moving_grid,
                             uint32_t max, uint32_t* count);

/* Fills `plan` with the cohort recipe: Rigid/Affine MI 32 bins, 25 % regular sampling,
 * 1000x500x250x100, shrink 8/4/2/1, sigmas 3/2/1/0 vox; SyN[0.1,3,0], CC radius 4,
 * 100x70x50x20, shrink 8/4/2/1, sigmas 3/2/1/0 vox (linear.rs:112-128, recipe.rs:93-112). */
void fw_plan_cohort(fw_plan* plan, int32_t channels);

/* One full registration of `channels` fixed/moving pairs (each fixed image on `fixed_grid`, each
 * moving on `moving_grid`, f64 host arrays) on the context's device. forward / inverse (3n each
 * on the fixed grid, may be NULL) receive the SyN fields (1Warp / 1InverseWarp). `cancel` may be
 * NULL; it is polled per level and per iteration. */
int fw_register(fw_ctx* ctx, const fw_plan* plan, const fw_grid* fixed_grid, const double* const* fixed,
                const fw_grid* moving_grid, const double* const* moving, double* forward, double* inverse,
                fw_result* result, const volatile int32_t* cancel);

/* The same registration from raw images already on the device (fw_upload_image handles, which it
 * does not consume): winsorising, windowing and the centre of mass run on the device, and the
 * SyN fields stay there (*forward / *inverse, new handles, may be NULL). No counterpart in the
 * reference, whose images cross PCIe per level: it serves callers that keep a cohort's volumes
 * resident in HBM (and the throughput measurement of the device path). */
int fw_register_resident(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, const fw_image* const* moving,
                         fw_field** forward, fw_field** inverse, fw_result* result, const volatile int32_t* cancel);

/* One SyN stage on images already on the device (crates/fastwarp-reg/src/syn.rs run_syn_from /
 * syn_fields with SynStart and the recipe's masks): the plan's SyN recipe (syn_* fields, channels,
 * channel weights and metrics) from `initial` (fixed points to moving points), the images taken
 * as they are (no preprocessing). fixed_mask / moving_mask (may be NULL; each on its own grid,
 * non-zero inside) limit the metrics as ITK resamples them onto the level grid every iteration
 * (syn.rs:505-580); prior / prior_inverse (both or neither) are the fields of the SyN stages
 * before this one, on the grid of this stage's last level: they enter every moving-side warp and
 * are collapsed with this stage's fields (syn.rs:746-774), so *forward / *inverse (new handles,
 * may be NULL) are the whole chain's. result (may be NULL) gets the per-level counts and
 * energies; fw_last_trace the iterations. */
int fw_syn_stage(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, const fw_image* const* moving,
                 const fw_affine* initial, const fw_image* fixed_mask, const fw_image* moving_mask, const fw_field* prior,
                 const fw_field* prior_inverse, fw_field** forward, fw_field** inverse, fw_result* result,
                 const volatile int32_t* cancel);

/* Many independent linear fits (centre of mass, Rigid, Affine Mattes MI on the fixed[i] /
 * moving[i] pair, each its own grids; fw_register_resident with run_syn 0): the reference runs
 * them one instance each (crates/fastwarp-cli/src/engine.rs:1016-1068). Here `workers` (1..16; 0
 * for 8) worker contexts of the device, each its own HIP stream and pool, take the fits in turn,
 * so their kernels and host round trips overlap on the device. results (may be NULL) holds one
 * entry per fit, each the bits of the fit's own run; an error of one fit fails the call. */
int fw_linear_batch(fw_ctx* ctx, const fw_plan* plan, uint32_t fits, const fw_image* const* fixed,
                    const fw_image* const* moving, uint32_t workers, fw_result* results, const volatile int32_t* cancel);

/* Several subjects registered to one template (the fixed images, `channels` of them, shared by
 * every subject): the reference runs a cohort as one registration per subject
 * (crates/fastwarp-cli/src/engine.rs:1016-1068, python/fastwarp/_api.py:1104-1137). Here the
 * template is preprocessed, its centre of mass taken and its smoothed levels built once; the
 * subjects' linear chains run side by side on `workers` worker contexts (as fw_linear_batch); then
 * the subjects' SyN stages run in lockstep, 8 at a time, each step of an iteration one launch over
 * all of them (CC channels only). moving holds subjects x channels handles (subject-major);
 * forward / inverse (may be NULL) and results (may be NULL) hold one entry per subject. Every
 * subject's result is the bits of its own fw_register_resident run; an error of one subject fails
 * the call. */
int fw_register_batch(fw_ctx* ctx, const fw_plan* plan, const fw_image* const* fixed, uint32_t subjects,
                      const fw_image* const* moving, uint32_t workers, fw_field** forward, fw_field** inverse,
                      fw_result* results, const volatile int32_t* cancel);
/* fw_last_trace for one subject (or fit) of the calling thread's last fw_register_batch or
 * fw_linear_batch (-1: no such subject). */
int32_t fw_last_batch_trace(uint32_t subject, fw_trace_step* out, int32_t capacity);

/* ---- Output files (crates/fastwarp-reg/src/transform/itk_mat.rs) ----
 * The affine of a registration as ANTs writes it (0GenericAffine.mat): an ITK MATLAB Level-4 file
 * of two f64 column vectors, `AffineTransform_double_3_3` (matrix row-major, then translation) and
 * `fixed` (the centre), 193 bytes (format_mat / write_mat_file, itk_mat.rs:150-190). Reading takes
 * the first transform of the file: the double or float forms of AffineTransform,
 * MatrixOffsetTransformBase, Rigid3DTransform and TranslationTransform (parse_mat, itk_mat.rs:102-133
 * with itk_params.rs:58-102). Host only, no device; on failure fw_last_error(NULL) gives the text. */
int fw_write_affine_mat(const char* path, const fw_affine* affine);
int fw_read_affine_mat(const char* path, fw_affine* out);

/* ITK HDF5 transform files, Composite.h5 and InverseComposite.h5 as antsRegistration writes them
 * (crates/fastwarp-reg/src/transform/itk_h5.rs:54-199 read_h5_file / format_h5 / write_h5_file,
 * on the HDF5 subset of transform/hdf5.rs): group TransformGroup/0 is the composite, members
 * 1, 2, ... the chain in queue order. An element is an affine (AffineTransform_double_3_3) or a
 * displacement field (DisplacementFieldTransform_double_3_3: its grid, and 3 values per voxel,
 * (x, y, z) displacements in mm with x fastest, as fw_field_download gives them; held and read as
 * f32 and written as f64, as the reference holds fields). Host only, no device; on failure
 * fw_last_error(NULL) gives the text. */
enum { FW_CHAIN_AFFINE = 0, FW_CHAIN_FIELD = 1 };
typedef struct {
  int32_t kind;     /* FW_CHAIN_AFFINE or FW_CHAIN_FIELD */
  fw_affine affine; /* an affine element */
  fw_grid grid;     /* a field element's grid */
  double* values;   /* a field element's 3 * voxels values: read when writing, filled when reading
                       (NULL: not filled) */
} fw_chain_element;
/* write_h5_file (itk_h5.rs:114-199): `count` >= 1 elements in queue order. */
int fw_write_composite_h5(const char* path, const fw_chain_element* elements, uint32_t count);
/* read_h5_file (itk_h5.rs:54-110): members in numeric order; the composite's own group gives no
 * element; affine, MatrixOffsetTransformBase, Rigid3D, Euler3D and translation classes become
 * affines (itk_params.rs:58-102). *count receives the chain's length; the first `capacity`
 * elements are filled, and a field element's values are copied where its `values` pointer (set
 * by the caller, 3 * voxels doubles) is not NULL: call once with capacity 0 for the length, once
 * to learn the grids, and once with the buffers. */
int fw_read_composite_h5(const char* path, fw_chain_element* elements, uint32_t capacity, uint32_t* count);
/* write_composite_pair with CompositePair::new / ::linear (ants_io.rs:100-141):
 * <prefix>Composite.h5 = [affine, field] and <prefix>InverseComposite.h5 = [inverse_field,
 * affine^-1]; with grid NULL (a linear-only registration) [affine] and [affine^-1]. The prefix is
 * used as given (a prefix ending in '/' names a folder). A singular affine is refused. */
int fw_write_composite_pair(const char* prefix, const fw_affine* affine, const fw_grid* grid, const double* field,
                            const double* inverse_field);
/* fw_write_composite_pair for fields resident on the device (fw_register_resident's outputs): both
 * are downloaded in f32 and written; neither handle is consumed. */
int fw_write_composite_fields(fw_ctx* ctx, const char* prefix, const fw_affine* affine, const fw_field* forward,
                              const fw_field* inverse);

/* ---- Serving: continuous batching of whole registrations on one device ------------------------
 * The reference runs several registrations on one GPU as independent device instances, as many
 * as its capacity model fits at once (crates/fastwarp-cli/src/engine.rs:1015-1068
 * `registrations_at_once`; python/fastwarp/_api.py:1104-1130 runs up to four at once on worker
 * threads, each on a backend from the idle pool, engine.rs:143-149). A server keeps
 * `slots` such instances (one fw_ctx each: its own HIP stream and buffer pool, on its own host
 * thread) and admits requests to them as slots and device memory free up: a request's footprint
 * is fw_plan_capacity's peak_bytes plus 256 MiB per instance (engine.rs:1000), admitted in
 * arrival order while it fits the budget beside the running ones. Requests overlap on the device,
 * and a request's result is the same bits alone and in any mix (no state is shared between slots
 * and every reduction has a fixed order).
 *   Images are f32 host arrays (x-fastest, one per channel) read by the request's slot when it
 * starts; they, and forward / inverse (3n interleaved f64 on the fixed grid, may be NULL), must
 * stay valid until fw_server_wait returns for the ticket. Every call is thread-safe except
 * fw_server_destroy, which cancels what has not ended, stops the slots and drops the tickets no
 * one waited for; no other call on the server may be in progress. Errors: the int status, and
 * fw_server_last_error() for the calling thread's last failure (a request's own message after
 * fw_server_wait). No counterpart in the reference's FFI: the entry points serve callers that
 * stream registrations through one GPU (a request loop), in place of one device per call. */
typedef struct fw_server fw_server;
typedef struct {
  double queued_s;          /* submission to start on a slot */
  double run_s;             /* start to end: uploads, registration, field downloads */
  double total_s;           /* submission to end (the request's latency) */
  uint64_t footprint_bytes; /* device bytes it was admitted with */
  int32_t slot;             /* the slot that ran it (-1: it never started) */
  int32_t pad;
} fw_served_timing;
/* `slots` 0 takes 4 (the reference's four at once, _api.py:1104; also the HIP hardware queues a
 * process gets by default), at most 16; a budget of
 * 0 takes fw_memory_budget's rule (free device memory less 1 GiB, or FASTWARP_GPU_MEMORY_BUDGET_MB). */
int fw_server_create(int32_t device, uint32_t slots, uint64_t budget_bytes, fw_server** out);
void fw_server_destroy(fw_server* server);
const char* fw_server_last_error(void);
/* Queues one registration (fw_register's arguments, f32 images); FW_ERR_INVALID for a bad plan,
 * grid or null image, or a footprint beyond the whole budget. */
int fw_server_submit(fw_server* server, const fw_plan* plan, const fw_grid* fixed_grid, const float* const* fixed,
                     const fw_grid* moving_grid, const float* const* moving, double* forward, double* inverse,
                     uint64_t* ticket);
/* Blocks until the request has ended; returns its status (FW_ERR_CANCELLED when cancelled) and
 * fills result and timing (either may be NULL). A ticket is waited for once. */
int fw_server_wait(fw_server* server, uint64_t ticket, fw_result* result, fw_served_timing* timing);
/* A waiting request ends at once; a running one at its next level or iteration. */
int fw_server_cancel(fw_server* server, uint64_t ticket);
/* Requests waiting and running, and the device bytes admitted to the running ones. */
int fw_server_counts(fw_server* server, uint32_t* waiting, uint32_t* running, uint64_t* bytes_in_use);

#ifdef __cplusplus
}
#endif

#endif /* FASTWARP_HIP_H */
