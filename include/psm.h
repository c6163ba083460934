/*
 * psm.h -- C ABI of the MI355X (gfx950) Kirchhoff post-stack 3D time
 * migration library, libpsm_hip.so.
 *
 * Plain C: pointers, sizes and status codes; no HIP or C++ types.  Every call
 * returns PSM_OK (0) or a nonzero PSM_ERR_* code; psm_last_error(ctx) then
 * holds a message.  A context owns one HIP stream and the device buffers its
 * calls need; calls on one context are serialised by an internal lock, calls
 * on different contexts may run concurrently (one context per host thread is
 * the intended use from OpenMP loops such as the reference's main.cpp:255).
 *
 * Reference interfaces replaced (sergeevsn/poststackmig3d):
 *   psm_kirchhoff_inline     MigrationKernel::kirchhoffKernel3DWithPadding,
 *                            include/migration_kernel.h:16-31,
 *                            src/migration_kernel.cpp:17-383
 *   psm_migrate_volume       the output-inline loop of src/main.cpp:254-383
 *                            together with the buffer staging it drives,
 *                            MigrationKernel::prepareInputBufferWithPadding
 *                            (migration_kernel.cpp:385-468) and
 *                            prepareVelocitySliceWithPadding (:470-511), and
 *                            the geometry of main.cpp:140-217
 *                            (output labels, estimateMaxVelocity :22-64,
 *                            inline aperture window :188-202, t0 table
 *                            :214-217)
 *   psm_survey_plan          main.cpp:140-217 alone (geometry only)
 */
#ifndef PSM_H
#define PSM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSM_OK 0
#define PSM_ERR_ARG 1     /* invalid argument (null pointer, bad size, ...) */
#define PSM_ERR_HIP 2     /* a HIP runtime call failed */
#define PSM_ERR_NOMEM 3   /* device or host allocation failed */
#define PSM_ERR_NODEV 4   /* no usable gfx950 device */

#define PSM_HOST 0        /* pointer arguments are host memory */
#define PSM_DEVICE 1      /* pointer arguments are device memory (HBM) */

/* Storage of the input traces in HBM (psm_set_storage).  Travel time,
 * interpolation weights and accumulation are float32 in every form. */
#define PSM_STORE_F32 0   /* float32 samples (the reference's arithmetic) */
#define PSM_STORE_F16 1   /* float16, each trace scaled by 1/max|x| */
#define PSM_STORE_BF16 2  /* bfloat16, each trace scaled by 1/max|x| */

typedef struct psm_ctx psm_ctx;

/* Library / context lifetime. */
const char* psm_version(void);
int psm_create(psm_ctx** ctx, int device);
void psm_destroy(psm_ctx* ctx);
const char* psm_last_error(const psm_ctx* ctx);
/* Use an existing HIP stream (hipStream_t passed as void*; NULL = the null
 * (default) stream) for every later call on this context.  A new context
 * uses a stream of its own. */
int psm_set_stream(psm_ctx* ctx, void* stream);
/* The HIP stream later calls on this context enqueue on (hipStream_t as
 * void*; NULL = the null stream).  The library orders nothing against other
 * streams: a caller that produces inputs on another stream makes this one
 * wait for that work first (an event, or psm_set_stream to that stream), and
 * waits on this one before reading the outputs of a PSM_DEVICE call. */
int psm_get_stream(psm_ctx* ctx, void** stream);
int psm_synchronize(psm_ctx* ctx);

/*
 * Per-inline drop-in for MigrationKernel::kirchhoffKernel3DWithPadding
 * (migration_kernel.h:16-31).  Same arguments in the same order, with the
 * std::vector arguments passed as pointer + length and the output
 * vector<vector<float>> [n_xl][n_t] passed as one contiguous host array.
 * xl_pos_indices is unused by the reference (migration_kernel.cpp:23) and is
 * accepted for signature parity (may be NULL).  All pointers are host
 * memory; the call is synchronous and fully overwrites output_slice.
 * Semantics are the reference's OpenMP build as main.cpp drives it: the
 * crosslines of one call are walked in order with per-sample parameter
 * caches that are not reset between crosslines (migration_kernel.cpp:64-152).
 */
int psm_kirchhoff_inline(psm_ctx* ctx,
                         const float* input_buffer,
                         size_t n_buf_il, size_t n_buf_xl, size_t n_t,
                         const float* velocity_slice,
                         float* output_slice, size_t n_xl,
                         const int32_t* il_pos_indices, size_t n_il_pos,
                         const int32_t* xl_pos_indices, size_t n_xl_pos,
                         int32_t current_il_pos, int32_t current_xl_start,
                         float dt, const float* t0_times,
                         float tan_theta_sq, float inline_bin_size,
                         float crossline_bin_size, int amp_correction);

/*
 * Survey description for the whole-volume path: the input grid (real traces
 * only, [n_il][n_xl][n_t], sorted labels) and the Config fields of
 * config_parser.h:7-20 that reach the kernel.
 */
typedef struct psm_survey {
    size_t n_il, n_xl, n_t;
    double dt;                       /* seconds (LookupTableResult::dt_seconds) */
    double inline_step;              /* metres */
    double crossline_step;           /* metres */
    int inline_padding;              /* traces per side */
    int crossline_padding;           /* traces per side */
    double angle_aperture;           /* degrees */
    int amp_correction;              /* 0/1 */
    const int32_t* inline_labels;    /* host [n_il], sorted; NULL = 1..n_il */
    const int32_t* crossline_labels; /* host [n_xl], sorted; NULL = 1..n_xl */
    float max_velocity;              /* estimateMaxVelocity (main.cpp:22-64);
                                        <= 0: probe the velocity cube */
} psm_survey;

/* Geometry derived from a survey (main.cpp:140-217).  Arrays are owned by the
 * plan and freed by psm_plan_free. */
typedef struct psm_plan {
    size_t n_out_il, n_out_xl, n_t;
    int max_aperture_il;             /* main.cpp:201 */
    float tan_theta_sq;              /* main.cpp:188-189 */
    float max_velocity;              /* main.cpp:192-195 */
    int32_t* out_inline_labels;      /* [n_out_il] */
    int32_t* out_crossline_labels;   /* [n_out_xl] */
    int32_t* out_to_in_il;           /* [n_out_il] real inline used (main.cpp:258-288) */
    int32_t* window;                 /* [2*n_out_il] data inline window [b0,b1) */
    float* t0_times;                 /* [n_t] */
} psm_plan;

/* vel_probe: the velocity cube (host or device per `where`) used when
 * survey->max_velocity <= 0; may be NULL otherwise.  A device cube is read
 * by a synchronous copy on the null stream: work on other streams that
 * writes it must be complete.  (The volume calls probe on their context's
 * stream instead.) */
int psm_survey_plan(const psm_survey* survey, const float* vel_probe, int where,
                    psm_plan* plan);
void psm_plan_free(psm_plan* plan);

/*
 * Whole-volume migration.  data and vel are [n_il][n_xl][n_t] float32 cubes
 * on the input grid (vel = Vrms of each image point, as the reference's
 * velocity provider hands it over per inline); out is
 * [n_out_il][n_out_xl][n_t] (psm_survey_plan gives the shape).  `where`
 * says whether the three pointers are host or device memory.  With
 * PSM_DEVICE the call enqueues on the context's stream and returns once the
 * work is queued; with PSM_HOST it copies, computes, copies back and
 * returns when out is valid.  Results equal main.cpp's loop over
 * kirchhoffKernel3DWithPadding on the same cube.
 */
int psm_migrate_volume(psm_ctx* ctx, const psm_survey* survey,
                       const float* data, const float* vel, float* out,
                       int where);

/*
 * psm_migrate_volume without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  data, vel and out are device memory; the call enqueues on
 * the context's stream and returns once the work is queued, and its result
 * equals psm_migrate_volume's with PSM_DEVICE.  The pair-layout cube, the
 * per-trace scales, the output-inline tasks and the t0 table live in a
 * static 64 MiB workspace that the library holds on every device it runs on
 * until the process ends: a survey that needs more than that (about
 * 8 B x (n_il + 1) x (n_xl + 1) x n_t with FP32 storage, half that with
 * FP16/BF16 storage, e.g. 101 x 101 x 501 but not 201 x 201 x 751 in FP32)
 * is rejected with PSM_ERR_ARG.  Every call on one device shares that
 * workspace: calls that overlap on one GPU (from any context or stream) race
 * and must be serialised by the caller.  Kernel timing
 * (psm_enable_timing) is not recorded by this call.
 */
int psm_migrate_volume_noalloc(psm_ctx* ctx, const psm_survey* survey,
                               const float* data, const float* vel, float* out);

/*
 * Adjoint of psm_migrate_volume (Kirchhoff demigration / modelling; no
 * reference counterpart): image [n_out_il][n_out_xl][n_t] -> data
 * [n_il][n_xl][n_t].  Walks the same pairs with the same sample index,
 * fraction, weight and mask as the forward operator and scatters
 * d[i] += w (1 - f) m, d[i+1] += w f m, so <M x, y> = <x, M^T y> up to float
 * rounding.  Crossline padding traces are zero inputs of the forward
 * operator and receive nothing.  Memory per `where`, as psm_migrate_volume.
 * The staged form keeps a row-pair buffer on the context of up to 8 GiB and
 * half the free device memory (one plane per output inline and row offset
 * of a pass; psm_set_tuning's rows_per_pass > 0 caps the row offsets per
 * pass); without room for one plane it adds each window into both data rows
 * directly.
 */
int psm_demigrate_volume(psm_ctx* ctx, const psm_survey* survey,
                         const float* image, const float* vel, float* data,
                         int where);

/*
 * Velocity scale (no reference counterpart): later volume calls on this
 * context use alpha * Vrms for travel times and weights and
 * alpha_mask * Vrms for the aperture mask, its crossline window and the
 * inline window (1, 1 = the reference operator).  psm_migrate_volume_dalpha
 * returns d(image)/d(alpha) at the current scales with the mask held at
 * alpha_mask, the tangent used for velocity-scale gradients.
 */
int psm_set_velocity_scale(psm_ctx* ctx, float alpha, float alpha_mask);
int psm_migrate_volume_dalpha(psm_ctx* ctx, const psm_survey* survey,
                              const float* data, const float* vel, float* out,
                              int where);

/*
 * Batched migration (no reference counterpart): several volumes -- e.g. a
 * velocity scan over scaled Vrms models and sub-volumes of any size -- in one
 * persistent launch.  Workgroups drain queues of (volume, t0 tile,
 * crossline group) items, latest tiles of all volumes first, so uneven
 * volumes balance; there is one queue per XCD (neighbouring items, which
 * share aperture rows in that XCD's L2), and a workgroup whose XCD's queue
 * is drained helps the others.  psm_batch_create packs nothing yet: it
 * plans every item and captures one HIP graph [queue reset, packing of each
 * distinct data cube, persistent kernel]; psm_batch_run replays it on the
 * context's stream (ms != NULL: waits and returns the elapsed time).
 * All item pointers are device memory and must stay valid while the batch
 * lives.  Item k's output equals psm_migrate_volume of that item at velocity
 * scale (alpha, alpha) by the same kernel form: the form follows the
 * context's psm_set_staging setting when the shape is set (create, set_shape,
 * autotune) -- staged by default (waves_per_block = output crosslines per
 * workgroup, 4 or 8, other values 8; crossline offsets per staged group from
 * psm_set_staging when staging is forced on, else 4), gather when staging is
 * switched off or a pair-layout row exceeds 1 GiB; the staged results do not
 * depend on the shape.
 * psm_batch_autotune times candidate launch shapes on the batch itself and
 * keeps the fastest: workgroup width x resident workgroups per CU, and for
 * the staged form also the offsets per staged group (2 or 4).
 * psm_batch_get_shape reports the shape in use (staged: 1 for the staged
 * form; steps: offsets per staged group, 0 for the gather form).
 * Storage: PSM_STORE_* for all items.
 */
typedef struct psm_batch psm_batch;
typedef struct psm_batch_item {
    psm_survey survey;  /* geometry and parameters of this volume */
    const float* data;  /* device [n_il][n_xl][n_t] */
    const float* vel;   /* device [n_il][n_xl][n_t] */
    float* out;         /* device [n_out_il][n_out_xl][n_t] */
    float alpha;        /* velocity scale of this volume (1 = vel as given) */
} psm_batch_item;
int psm_batch_create(psm_ctx* ctx, const psm_batch_item* items, int n_items, int storage,
                     psm_batch** batch);
int psm_batch_set_shape(psm_batch* batch, int waves_per_block, int blocks_per_cu);
int psm_batch_autotune(psm_batch* batch, int* waves_per_block, int* blocks_per_cu,
                       float* best_ms);
int psm_batch_get_shape(psm_batch* batch, int* staged, int* waves_per_block,
                        int* blocks_per_cu, int* steps);
int psm_batch_run(psm_batch* batch, float* ms);
void psm_batch_destroy(psm_batch* batch);

/* sum x[i] y[i] accumulated in float64 on the GPU (deterministic order);
 * x, y host or device per `where`. */
int psm_dot_f64(psm_ctx* ctx, const float* x, const float* y, size_t n, int where,
                double* result);

/* Pair counts of the work unit for a survey (no reference counterpart):
 * p_ap = (output sample, real aperture trace) pairs inside the aperture
 * mask, p_g = those that also gather (i < n_t - 1).  ap_image / g_image
 * (nullable, [n_out_il][n_out_xl][n_t]) receive the per-sample counts.
 * vel and the images are host or device memory per `where`. */
int psm_count_pairs(psm_ctx* ctx, const psm_survey* survey, const float* vel,
                    int where, double* p_ap, double* p_g,
                    float* ap_image, float* g_image);

/* Launch shape of the migration kernel (no reference counterpart):
 * waves_per_block = output crosslines per workgroup (1, 2, 4, 8 or 16);
 * rows_per_pass = aperture row offsets summed per launch (0 = the whole
 * aperture in one launch; smaller values bound each launch's cache working
 * set to a few aperture rows, later launches accumulate; -1, the default =
 * automatic: one launch for the staged form and for FP16/BF16 gathers, 8
 * for FP32 gathers; the staged adjoint: as many row offsets per pass as its
 * 8 GiB row-pair buffer holds).  Results do not depend on the shape beyond
 * FP32 summation order. */
int psm_set_tuning(psm_ctx* ctx, int waves_per_block, int rows_per_pass);
int psm_get_tuning(psm_ctx* ctx, int* waves_per_block, int* rows_per_pass);

/* Forward-migration kernel form (no reference counterpart): on = 1 stages
 * each aperture trace's time window in LDS once per workgroup of output
 * crosslines (waves_per_block 4 or 8) and gathers from there; on = 0 gathers
 * every sample through the vector memory path; on = -1 (the default) stages
 * with 8 output crosslines per workgroup and 4 crossline offsets per staged
 * group in one launch.  steps (2 or 4) = crossline offsets per staged group
 * when on = 1.  The two rows +db, -db of each row pair are staged as one
 * window of summed FP32 pairs (they share every step's travel time and
 * weight); FP16/BF16 pairs are widened and multiplied by their trace's scale
 * as they are staged, so the staged steps are FP32 for every storage form.
 * The staged form
 * applies to whole-volume calls and per-inline calls with contiguous rows;
 * other calls use the gather form.  The switch also selects the adjoint's
 * form (psm_demigrate_volume): staged (LDS accumulator windows per column,
 * 8 output crosslines per workgroup, each window added once into a
 * row-pair buffer that a second kernel adds into both data rows of the
 * pair) unless on = 0.
 * Results do not depend on the form beyond FP32 summation order. */
int psm_set_staging(psm_ctx* ctx, int on, int steps);
int psm_get_staging(psm_ctx* ctx, int* on, int* steps);

/* Trace storage used by later calls on this context (no reference
 * counterpart): PSM_STORE_F32 (default), PSM_STORE_F16 or PSM_STORE_BF16.
 * The reduced forms halve the bytes each gather reads; results then differ
 * from the float32 form by the rounding of the stored samples. */
int psm_set_storage(psm_ctx* ctx, int storage);

/* Kernel-level timing (no reference counterpart): when enabled, each
 * psm_migrate_volume records HIP events on the context's stream around the
 * trace packing kernel and the migration kernel; psm_last_timing waits for
 * the last call and returns the two durations in milliseconds. */
typedef struct psm_timing {
    float pack_ms;
    float migrate_ms;
} psm_timing;
int psm_enable_timing(psm_ctx* ctx, int on);
int psm_last_timing(psm_ctx* ctx, psm_timing* timing);

/* Table velocity provider on the GPU.  Replaces
 * TableVelocityProvider::interpolateVelocity / getVelocity
 * (velocity_reader.cpp:320-370) evaluated for every sample of the data grid:
 * vel[il][xl][t] at (il_labels[il], crossline_labels[xl], t * dt) is the
 * first point at 4-D distance < 1e-10 (distance over inline and crossline
 * label differences and time, in double), else the mean of the points'
 * velocities with float weights 1 / (1 + d) accumulated in point order, and
 * the smallest point velocity where that mean is not positive.  `points` is
 * host memory; `vel` is host or device memory per `where`. */
typedef struct psm_velocity_point {
    int32_t inline_label;
    int32_t crossline_label;
    double time;
    float velocity;
} psm_velocity_point;
int psm_table_velocity(psm_ctx* ctx, const psm_velocity_point* points, size_t n_points,
                       const int32_t* inline_labels, size_t n_il,
                       const int32_t* crossline_labels, size_t n_xl,
                       size_t n_t, double dt, float* vel, int where);

/* Page-locked host memory for whole-volume ingest (no reference
 * counterpart): a cube read once into such a buffer reaches the device in one
 * DMA transfer at full link rate.  psm_free_host(NULL) is a no-op. */
int psm_alloc_host(size_t bytes, void** ptr);
void psm_free_host(void* ptr);

#ifdef __cplusplus
}
#endif

#endif /* PSM_H */
