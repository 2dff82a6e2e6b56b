/*
 * r3dg.h -- C ABI of the MI355X (gfx950) Census-SGM stereo matcher.
 *
 * Every entry point is `extern "C"`, takes plain pointers, sizes and POD
 * descriptors, never throws, and returns an r3dg_status.  The message of the
 * last failure of a context is read with r3dg_last_error(); failures that
 * happen before a context exists go to r3dg_global_error().
 *
 * Images are strided 2-D views (r3dg_image).  `memory` says whether `data`
 * is host memory (the library copies it to and from the device) or device
 * memory of the context's GPU (used in place, no copy).  Outputs follow the
 * same rule.  All calls of one context run on that context's HIP stream
 * and return after the results are complete, like the reference API they
 * replace; several host threads may share a context (calls are serialised by
 * a per-context mutex).  Device inputs must be complete, and device outputs
 * no longer in use by other work, when a call starts: the library does not
 * order its stream after the caller's streams (the Python binding
 * synchronises torch's current stream before it passes a tensor).
 *
 * Semantics follow the reference matcher (rhiestan/Regard3D_MVS,
 * src/r3d/stereo/stereo_matcher.hpp); each declaration cites the reference
 * interface it replaces.
 */
#ifndef R3DG_H_
#define R3DG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define R3DG_ABI_VERSION 3

typedef enum r3dg_status {
  R3DG_OK = 0,
  R3DG_ERROR_INVALID_ARGUMENT = 1, /* maps to r3d::Error (stereo_matcher.cpp:81-111, 128-148) */
  R3DG_ERROR_OUT_OF_MEMORY = 2,    /* device or pinned allocation failed */
  R3DG_ERROR_DEVICE = 3,           /* a HIP runtime call failed */
  R3DG_ERROR_CANCELLED = 4,        /* maps to r3d::Cancelled (core/cancel.hpp:10-13) */
  R3DG_ERROR_BUDGET = 5,           /* request larger than the context's HBM budget */
  R3DG_ERROR_NOT_FOUND = 6,        /* unknown ticket or view id */
  R3DG_ERROR_INTERNAL = 7,
  R3DG_PENDING = 8                 /* a submitted pair that has not finished yet */
} r3dg_status;

typedef enum r3dg_memory { R3DG_HOST = 0, R3DG_DEVICE = 1 } r3dg_memory;

/* Element format of an image view.  R3DG_FORMAT_U8X257 marks uint8 data given
 * where a uint16 image is expected (left, right and view images, Census and
 * cost-volume inputs): the bytes are transferred as they are and widened on
 * the device to u8 * 257, the reference's 8-to-16-bit convention
 * (io/image_io.hpp:25-26); stride is then in bytes. */
typedef enum r3dg_format { R3DG_FORMAT_NATIVE = 0, R3DG_FORMAT_U8X257 = 1 } r3dg_format;

/* A strided single-channel image; stride in elements (>= width).  A view with
 * data == NULL or width == 0 or height == 0 is empty (masks: all valid).
 * Host memory from r3dg_host_alloc is transferred by DMA directly; other
 * host memory is staged through the context's pinned blocks.  Device memory
 * must hold its final contents when the call starts: the library's stream
 * does not wait for work the caller queued on other streams. */
typedef struct r3dg_image {
  void* data;
  int32_t width;
  int32_t height;
  int64_t stride;
  int32_t memory; /* r3dg_memory */
  int32_t format; /* r3dg_format */
} r3dg_image;

typedef enum r3dg_subpixel {
  R3DG_SUBPIXEL_NONE = 0,
  R3DG_SUBPIXEL_PARABOLA = 1,
  R3DG_SUBPIXEL_EQUIANGULAR = 2
} r3dg_subpixel;

/* Mirror of r3d::StereoOptions (stereo_matcher.hpp:30-57); the reference's
 * `scalar_kernels` switch selects CPU kernels and has no counterpart here. */
typedef struct r3dg_options {
  int32_t p1;
  int32_t p2;
  int32_t p2_adapt;
  int32_t invalid_cost;
  int32_t subpixel; /* r3dg_subpixel */
  float uniqueness;
  int32_t lr_max_diff;
  int32_t lr_full_match;
  int32_t speckle_size;
  float speckle_diff;
  int32_t median;
  int32_t levels;
  int32_t range_radius;
  int32_t range_margin;
  int32_t range_fill;
  int32_t tile_size;
  int32_t tile_margin;
  int32_t tile_blend;
  uint64_t memory_budget;
} r3dg_options;

/* Stages timed with HIP events when profiling is on (r3dg_set_profiling). */
enum {
  R3DG_STAGE_PYRAMID = 0,   /* pyramid_down of images and masks */
  R3DG_STAGE_PREPARE = 1,   /* Census, valid intervals, matching costs */
  R3DG_STAGE_RANGES = 2,    /* ranges_from_coarse */
  R3DG_STAGE_LAYOUT = 3,    /* tile label layout (including its read-back) */
  R3DG_STAGE_AGGREGATE = 4, /* the 8 SGM path kernels */
  R3DG_STAGE_SELECT = 5,    /* WTA, sub-pixel, checks */
  R3DG_STAGE_BLEND = 6,     /* tile blending */
  R3DG_STAGE_FILTER = 7,    /* left-right check, speckle and median filters */
  R3DG_STAGES = 8
};

/* Per-call statistics (DisparityMap fields, stereo_matcher.hpp:61-68, plus timings). */
typedef struct r3dg_match_info {
  int32_t disparity_min;   /* DisparityMap::disparity_min */
  int32_t num_disparities; /* DisparityMap::num_disparities (aligned to 16) */
  int32_t levels;          /* DisparityMap::levels */
  int32_t tiles_level0;    /* tiles of the finest level, one direction */
  uint64_t cells;          /* DisparityMap::cells: aggregated cells, all levels and directions */
  uint64_t tile_pixels;    /* sum of tile pixels over levels and directions */
  double seconds;          /* wall time of the call */
  uint64_t aggregate_launches; /* path-kernel launches */
  uint64_t aggregate_bytes;    /* algorithmic bytes of the path kernels: sum of (40 - 2 w) C + 80 P,
                                  w the streams the 8 scans of a chunk ran on (38 C + 80 P for one) */
  double stage_seconds[R3DG_STAGES]; /* device time per stage (0 unless profiling) */
} r3dg_match_info;

/* Optional integer outputs of r3dg_match at level 0 (views may be empty).
 * `labels`: int16 WTA disparity (disparity_min + best label) of the tile that
 * dominates the blend of each pixel, INT16_MIN where the blend is invalid.
 * `lr_mask`: uint8, 1 where the left-right check invalidated a valid pixel. */
typedef struct r3dg_match_debug {
  r3dg_image labels;
  r3dg_image lr_mask;
} r3dg_match_debug;

/* One pair of r3dg_match_batch: inputs as for r3dg_match, outputs written by
 * the call (status: R3DG_OK or this pair's error; info.seconds and
 * info.stage_seconds are those of the whole batch). */
typedef struct r3dg_pair {
  r3dg_image left, right;             /* uint16 */
  r3dg_image left_valid, right_valid; /* uint8, empty views: all valid */
  int32_t disparity_min;
  int32_t num_disparities;
  r3dg_image disparity;               /* float, left size */
  r3dg_image confidence;              /* float, left size, may be empty */
  r3dg_match_info info;
  int32_t status;
  int32_t reserved;
} r3dg_pair;

typedef struct r3dg_ctx r3dg_ctx;

/* ---- library and context ------------------------------------------------ */

int32_t r3dg_abi_version(void);

/* Defaults of r3d::StereoOptions (stereo_matcher.hpp:30-57). */
void r3dg_default_options(r3dg_options* options);

/* Creates a context on HIP device `device`; `hbm_budget` bytes (0: 90% of the
 * free memory) bound the device memory the context may hold.  No reference
 * counterpart: the reference runs on the process-wide CPU thread pool
 * (core/thread_pool.hpp:65-77) with the per-call MemoryGate
 * (stereo_matcher.cpp:580-601). */
int32_t r3dg_ctx_create(int32_t device, uint64_t hbm_budget, r3dg_ctx** out);
int32_t r3dg_ctx_destroy(r3dg_ctx* ctx);

/* Page-locked host memory from the library's HIP runtime (no reference
 * counterpart): images in it move between host and device by DMA with no
 * staging copy.  r3dg_host_free releases it. */
int32_t r3dg_host_alloc(size_t bytes, void** out);
int32_t r3dg_host_free(void* p);

/* Copies the message of the context's last failure into buf (NUL-terminated)
 * and returns the status of that failure. */
int32_t r3dg_last_error(const r3dg_ctx* ctx, char* buf, size_t size);
int32_t r3dg_global_error(char* buf, size_t size);

/* Records HIP events around every stage of later calls on the context's
 * stream and reports their durations in r3dg_match_info::stage_seconds. */
int32_t r3dg_set_profiling(r3dg_ctx* ctx, int32_t enabled);

/* Test hook, no counterpart in the reference: after `after` further device
 * allocations of this context succeed, the next one fails once as if the
 * device were out of memory (R3DG_ERROR_OUT_OF_MEMORY from the call that
 * makes it); a negative value disarms it.  Lets a host test force the
 * allocation-failure path of every call and of the request loop. */
int32_t r3dg_inject_alloc_failure(r3dg_ctx* ctx, int64_t after);

/* Cancellation flag polled between pyramid levels and tile stages; replaces
 * CancellationToken polling in parallel_for (core/thread_pool.cpp:41-70).
 * The flag stays set until it is cleared (requested = 0), as a token stays
 * requested until reset (core/cancel.hpp:22-24). */
int32_t r3dg_cancel(r3dg_ctx* ctx, int32_t requested);

/* A caller's cancellation test, polled at the same points as the flag: a
 * non-zero poll(user) cancels the running call (R3DG_ERROR_CANCELLED).  NULL
 * removes it.  Replaces current_cancellation() (core/cancel.hpp:39-44): the
 * C++ layer passes the calling thread's CancellationToken. */
int32_t r3dg_set_cancel_poll(r3dg_ctx* ctx, int32_t (*poll)(void* user), void* user);

/* ---- matcher stages ------------------------------------------------------ */

/* census_transform (stereo/census.hpp:20): 9x7 Census, 62 bits, first
 * neighbour in bit 61, replicated borders.  img: uint16; out: uint64 image of
 * the same size. */
int32_t r3dg_census(r3dg_ctx* ctx, r3dg_image img, r3dg_image out);

/* pyramid_down<uint16_t/uint8_t> (core/pyramid.hpp:20): [1 4 6 4 1]/16
 * separable decimation, (sum + 128) >> 8; out is ceil(w/2) x ceil(h/2). */
int32_t r3dg_pyramid_down_u16(r3dg_ctx* ctx, r3dg_image img, r3dg_image out);
int32_t r3dg_pyramid_down_u8(r3dg_ctx* ctx, r3dg_image img, r3dg_image out);

/* detail::sgm_cost_volume (stereo_matcher.hpp:98-101): the whole image as one
 * tile; S of every pixel's labels in row-major pixel order.  lo/num are uint16
 * images of the left size (both empty: full range).  out_cells must equal the
 * sum of num (w * h * align16(num_disparities) for the full range). */
int32_t r3dg_cost_volume(r3dg_ctx* ctx, r3dg_image left, r3dg_image right, int32_t disparity_min,
                         int32_t num_disparities, const r3dg_options* options, r3dg_image lo, r3dg_image num,
                         uint16_t* out, uint64_t out_cells, int32_t out_memory);

/* match_stereo (stereo_matcher.hpp:82-85).  left/right: uint16; masks: uint8
 * (255 valid, empty view: all valid); disparity/confidence: float images of
 * the left size (confidence may be empty); debug may be NULL. */
int32_t r3dg_match(r3dg_ctx* ctx, r3dg_image left, r3dg_image right, r3dg_image left_valid,
                   r3dg_image right_valid, int32_t disparity_min, int32_t num_disparities,
                   const r3dg_options* options, r3dg_image disparity, r3dg_image confidence,
                   const r3dg_match_debug* debug, r3dg_match_info* info);

/* match_stereo for several pairs in one call; no counterpart in the reference,
 * whose callers loop over pairs (mvs/multi_stereo.cpp:264-293 calls
 * match_stereo once per partner).  The levels of all pairs run in lockstep,
 * finest levels together, and the jobs of a level share the cost, path and
 * selection launches; pairs are grouped so that a group fits the context's
 * memory budget.  Every pair's outputs are bit-identical to r3dg_match on it
 * alone.  Pairs that fail validation get their status and are skipped; the
 * call returns the first pair error, or R3DG_OK. */
int32_t r3dg_match_batch(r3dg_ctx* ctx, r3dg_pair* pairs, int32_t count, const r3dg_options* options);

/* ---- serving (no counterpart in the reference; for callers that match
 * pairs as they arrive) ---------------------------------------------------
 * r3dg_server_start runs a request loop on its own thread.  Submitted pairs
 * wait until their estimated device footprint fits hbm_budget (0: 80% of the
 * context budget) next to the pairs in flight; every step of the loop then
 * matches the next level of every admitted pair with shared launches, so new
 * pairs join while others are on the GPU (continuous batching).  A finished
 * pair's outputs are downloaded on the loop's own copy stream while its next
 * steps run and written to host buffers by the loop's copier thread; then its
 * ticket completes; pair->info.seconds is the time from submission to
 * completion.  A pair and the images it points to
 * must stay valid until its ticket completes.  Each pair's outputs equal
 * those of r3dg_match on it alone.  Direct calls on the context interleave
 * with the loop's steps. */
int32_t r3dg_server_start(r3dg_ctx* ctx, const r3dg_options* options, uint64_t hbm_budget);
/* Finishes the submitted pairs, then stops the loop. */
int32_t r3dg_server_stop(r3dg_ctx* ctx);
/* Counters of the running loop (since r3dg_server_start).  A pair whose
 * start or step runs out of device memory while other pairs are in flight is
 * put back at the head of the queue (requeued); the loop then admits nothing
 * new until a pair completes and lowers its admission budget by a quarter
 * (admission_scale) after a step that ran out.  Running out is the context's
 * budget (R3DG_ERROR_BUDGET) or a failed device allocation
 * (R3DG_ERROR_OUT_OF_MEMORY); a pair alone in flight fails with it. */
typedef struct r3dg_server_stats {
  uint64_t submitted;        /* tickets issued */
  uint64_t completed;        /* pairs finished, with any status */
  uint64_t failed;           /* of those, finished with an error */
  uint64_t steps;            /* loop steps (one level of every admitted pair) */
  uint64_t requeued;         /* pairs put back after running out of memory */
  uint64_t pending;          /* waiting for admission now */
  uint64_t active;           /* admitted and not finished now */
  uint64_t in_use;           /* estimated device bytes of the admitted pairs */
  uint64_t view_bytes;       /* device bytes of the cached views */
  double admission_scale;    /* factor on the admission budget, 1 at start */
  double loop_seconds;       /* loop thread: time in its steps (admission, launches, waits) */
  double wait_seconds;       /* of that, time waiting for the device (stream synchronisations) */
} r3dg_server_stats;
int32_t r3dg_server_stats_get(r3dg_ctx* ctx, r3dg_server_stats* out);
/* Test hook, no counterpart in the reference: in the next loop step with at
 * least two pairs in flight, the device allocation after `after` more fails
 * once with R3DG_ERROR_OUT_OF_MEMORY (dropped if the step makes fewer);
 * negative disarms it.  Forces the requeue path from the host. */
int32_t r3dg_server_inject_step_failure(r3dg_ctx* ctx, int64_t after);
/* left_view / right_view: ids given to r3dg_view_put, used in place of the
 * pair's left + left_valid (right + right_valid) images when non-zero. */
int32_t r3dg_submit(r3dg_ctx* ctx, r3dg_pair* pair, uint64_t left_view, uint64_t right_view, uint64_t* ticket);
/* R3DG_PENDING while the pair runs, else its status (the ticket is released). */
int32_t r3dg_poll(r3dg_ctx* ctx, uint64_t ticket);
/* Waits up to timeout_seconds (< 0: no limit); returns as r3dg_poll. */
int32_t r3dg_wait(r3dg_ctx* ctx, uint64_t ticket, double timeout_seconds);
/* Keeps an image (uint16) and its mask (uint8, may be empty) on the device
 * under a non-zero id, with its pyramids and Census built on first use and
 * reused by every pair that names the id.  Replaces an existing
 * view of that id.  While a request loop runs, host images are copied by the
 * call and uploaded by the loop before its next step. */
int32_t r3dg_view_put(r3dg_ctx* ctx, uint64_t view, r3dg_image image, r3dg_image valid);
/* Drops a view; pairs in flight that use it keep it until they finish. */
int32_t r3dg_view_evict(r3dg_ctx* ctx, uint64_t view);

/* remove_speckles (stereo_matcher.hpp:89), in place on a float image. */
int32_t r3dg_remove_speckles(r3dg_ctx* ctx, r3dg_image disparity, int32_t min_size, float max_diff);

/* The speckle filter of r3dg_remove_speckles, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The map is filtered in place and must be device memory,
 * contiguous (stride == width), of at most R3DG_STATIC_SPECKLE_PIXELS pixels;
 * other maps are R3DG_ERROR_INVALID_ARGUMENT.  The union-find scratch is a
 * fixed 32 MiB device array per GPU, held until the process ends and shared
 * by every call on that GPU: calls through this library take turns and return
 * after their kernels finish, but work that other code queues on the same
 * array at the same time would race.  Results equal r3dg_remove_speckles. */
#define R3DG_STATIC_SPECKLE_PIXELS 4194304
int32_t r3dg_remove_speckles_static(r3dg_ctx* ctx, r3dg_image disparity, int32_t min_size, float max_diff);

/* median_filter_3x3 (stereo_matcher.hpp:91). */
int32_t r3dg_median_filter_3x3(r3dg_ctx* ctx, r3dg_image disparity, r3dg_image out);

/* detail::ranges_from_coarse (stereo_matcher.hpp:121-125); lo/num: uint16
 * images of width x height. */
int32_t r3dg_ranges_from_coarse(r3dg_ctx* ctx, r3dg_image coarse, int32_t width, int32_t height,
                                r3dg_image fine_valid, r3dg_image partner_valid, int32_t partner_width,
                                int32_t partner_sign, int32_t disparity_min, int32_t num_disparities,
                                const r3dg_options* options, int32_t mirrored, r3dg_image lo, r3dg_image num);

/* detail::automatic_levels (stereo_matcher.hpp:128). Host-only arithmetic. */
int32_t r3dg_automatic_levels(int32_t width, int32_t height, int32_t num_disparities);

/* ---- row kernels (for callers that drive the row loop themselves) --------
 * The reference's row-kernel interface (stereo/census.hpp:25-29,
 * stereo/sgm_kernels.hpp:16-97) on the GPU: every pointer is host memory
 * laid out as the reference describes; the call uploads what the kernel
 * reads, runs it on the context's stream and writes the reference's outputs
 * back (only the elements the reference writes). */

/* census_row: rows[j] points to row y - 3 + j of an image padded by 4
 * replicated pixels on both sides, at the padded start (n + 8 elements);
 * writes out[0 .. n). */
int32_t r3dg_census_row(r3dg_ctx* ctx, const uint16_t* const rows[7], uint64_t* out, size_t n);

/* sgm::RowLabels (sgm_kernels.hpp:28-39): pixel x has labels lo[x] ..
 * lo[x] + num[x] - 1 (num a positive multiple of 16), packed from offset[x]
 * (offset[width] = total); path buffers pad every pixel by 8 on both sides. */
typedef struct r3dg_row_labels {
  const int32_t* lo;
  const uint16_t* num;
  const size_t* offset; /* [width + 1] */
} r3dg_row_labels;

/* sgm::census_costs (sgm_kernels.hpp:41-44): cost[offset[x] + j] =
 * popcount(left[x] ^ rev[width - 1 - x + lo[x] + j - rev_origin]). */
int32_t r3dg_census_costs(r3dg_ctx* ctx, const uint64_t* left, const uint64_t* rev, size_t width,
                          const r3dg_row_labels* labels, int32_t rev_origin, uint16_t* cost);

/* sgm::SgmRowArgs (sgm_kernels.hpp:54-73), same fields and meaning.  scratch:
 * only its first buffer (the zero start of every path, path_stride(max_num)
 * elements) is read; the reference's other scratch buffers are its own
 * working storage and are not written back. */
typedef struct r3dg_sgm_row_args {
  const uint16_t* cost;
  const uint16_t* p2;       /* [4][width] */
  const uint16_t* prev;     /* [3][prev_labels path size] */
  const uint16_t* prev_min; /* [3][width] */
  uint16_t* cur;            /* [3][labels path size] */
  uint16_t* cur_min;        /* [3][width] */
  uint16_t* scratch;
  uint16_t* sum;
  r3dg_row_labels labels;
  r3dg_row_labels prev_labels;
  size_t width;
  size_t max_num;
  uint16_t p1;
  int32_t direction;
  int32_t has_prev;
  int32_t accumulate;
} r3dg_sgm_row_args;

/* sgm::sgm_row (sgm_kernels.hpp:46-76): one row of one 4-path SGM pass. */
int32_t r3dg_sgm_row(r3dg_ctx* ctx, const r3dg_sgm_row_args* args);

/* sgm::wta_row (sgm_kernels.hpp:78-88): best (absolute label), best cost and
 * second best per pixel; right_cost / right_best (may be NULL) are updated for
 * every i = x - lo[x] - k + right_shift as the reference does. */
int32_t r3dg_wta_row(r3dg_ctx* ctx, const uint16_t* sum, size_t width, const r3dg_row_labels* labels, uint16_t* best,
                     uint16_t* best_cost, uint16_t* second, uint16_t* right_cost, uint16_t* right_best,
                     ptrdiff_t right_shift);

/* ---- rectification (the step before the matcher) ------------------------- */

/* Camera models of core/camera_model.hpp:13-20 (OpenMVG formulas; params in
 * that order: radial_k1 k1; radial_k3 k1 k2 k3; brown_t2 k1 k2 k3 t1 t2;
 * fisheye k1..k4). */
typedef enum r3dg_camera_model {
  R3DG_CAMERA_PINHOLE = 0,
  R3DG_CAMERA_RADIAL_K1 = 1,
  R3DG_CAMERA_RADIAL_K3 = 2,
  R3DG_CAMERA_BROWN_T2 = 3,
  R3DG_CAMERA_FISHEYE = 4
} r3dg_camera_model;

/* CameraIntrinsics (core/camera_model.hpp:32-41) with its Pose
 * (core/camera.hpp:9-12): rotation row-major, world to camera axes; centre
 * in world coordinates.  A PinholeCamera (core/camera.hpp:58-61) is the same
 * record with model R3DG_CAMERA_PINHOLE. */
typedef struct r3dg_camera {
  int32_t model; /* r3dg_camera_model */
  int32_t width;
  int32_t height;
  int32_t reserved;
  double fx, fy, cx, cy;
  double params[5];
  double rotation[9];
  double center[3];
} r3dg_camera;

/* Element types of r3dg_rectify. */
typedef enum r3dg_dtype { R3DG_U8 = 0, R3DG_U16 = 1, R3DG_F32 = 2 } r3dg_dtype;

/* rectify_image<T> (rectify/rectification.hpp:80-82, .cpp:288-350), one
 * channel: resamples `src`, an image of `camera`, into the pinhole camera
 * `target` of the same centre; bicubic (Catmull-Rom, core/sampling.hpp:21-55)
 * from the pyramid level whose resolution is closest above 1/1.5 of the
 * target's; out-of-image pixels are 0.  out: target width x height of dtype;
 * valid (uint8, may be empty) receives 255 inside and 0 outside.  Errors as
 * the reference: a source whose size differs from the camera's, or a target
 * with another centre, is R3DG_ERROR_INVALID_ARGUMENT. */
int32_t r3dg_rectify(r3dg_ctx* ctx, int32_t dtype, r3dg_image src, const r3dg_camera* camera,
                     const r3dg_camera* target, r3dg_image out, r3dg_image valid);

/* ---- combining the pairs of a view (the step after the matcher) ---------- */

/* Per pair of a view: the PairGeometry of mvs/multi_stereo.cpp:26-33, as the
 * caller computes it from its RectifiedPair and the grid camera
 * (multi_stereo.cpp:84-96): grid pixel -> rectified pixel homography
 * K_r R_r R_g^T K_g^-1 (row-major), f * baseline, disparity offset, the
 * partner camera's centre; and the pair's maps (float, rectified left image;
 * confidence may be empty). */
typedef struct r3dg_pair_geometry {
  double grid_to_rect[9];
  double fb;
  double offset;
  double partner_center[3];
  r3dg_image disparity;
  r3dg_image confidence;
} r3dg_pair_geometry;

/* The grid camera of the depth map: size, K_g^-1 and R_g^T (row-major) and
 * its centre (multi_stereo.cpp:84-85, 193-195). */
typedef struct r3dg_grid {
  int32_t width;
  int32_t height;
  double inv_k[9];
  double rot_t[9];
  double center[3];
} r3dg_grid;

/* MultiStereoOptions (mvs/multi_stereo.hpp:25-30), with cos(min_angle_deg)
 * precomputed by the caller (multi_stereo.cpp:98). */
typedef struct r3dg_multi_options {
  double inlier_px;
  double max_reprojection_px;
  double cos_min_angle;
  int32_t min_pairs;
  int32_t reserved;
} r3dg_multi_options;

/* combine_pairs (mvs/multi_stereo.hpp:43-44, .cpp:74-216): per grid pixel,
 * the hypothesis with the most agreeing pairs, refined by weighted least
 * squares on inverse depth, kept if it has a majority, min_pairs pairs and a
 * wide enough triangulation angle.  depth, confidence: float; pair_mask:
 * uint16 (bit i: pair i used); pixels without a result are 0.  At most 16
 * pairs (R3DG_ERROR_INVALID_ARGUMENT beyond, as the reference). */
int32_t r3dg_combine_pairs(r3dg_ctx* ctx, const r3dg_grid* grid, const r3dg_pair_geometry* pairs, int32_t count,
                           const r3dg_multi_options* options, r3dg_image depth, r3dg_image confidence,
                           r3dg_image pair_mask);

/* ---- file formats (host only; no context, errors via r3dg_global_error) ---- */

/* read_pfm / write_pfm (io/pfm.hpp, io/pfm.cpp:30-102): "Pf" (1 channel) or
 * "PF" (3, interleaved), rows stored bottom to top, little-endian when the
 * scale is negative (written as -1).  r3dg_pfm_header reads the size; out
 * receives height rows of width * channels floats, `stride` floats apart. */
int32_t r3dg_pfm_header(const char* path, int32_t* width, int32_t* height, int32_t* channels);
int32_t r3dg_read_pfm(const char* path, float* out, int64_t stride);
int32_t r3dg_write_pfm(const char* path, const float* data, int32_t width, int32_t height, int32_t channels,
                       int64_t stride);

/* The .r3dd depth map file of write_depth_map / read_depth_map
 * (mvs/depth_map.hpp:59-62, depth_map.cpp:94-221), version 2: view, the
 * grid camera (intrinsics, row-major rotation, centre), up to 16 partner
 * views, then depth (float), confidence (float), pair mask (uint16), support
 * (uint8) and colour (uint8 RGB) of width x height, rows contiguous. */
typedef struct r3dg_depth_map {
  uint32_t view;
  int32_t camera_width;
  int32_t camera_height;
  uint32_t num_partners;
  double fx, fy, cx, cy;
  double rotation[9];
  double center[3];
  uint32_t partners[16];
  int32_t width;
  int32_t height;
  float* depth;
  float* confidence;
  uint16_t* pair_mask;
  uint8_t* support;
  uint8_t* color;
} r3dg_depth_map;

/* Reads everything but the images (the caller then sizes the buffers). */
int32_t r3dg_read_depth_map_header(const char* path, r3dg_depth_map* map);
/* Reads the whole file into `map`, whose image pointers must hold
 * width * height elements (colour: 3 per pixel). */
int32_t r3dg_read_depth_map(const char* path, r3dg_depth_map* map);
int32_t r3dg_write_depth_map(const char* path, const r3dg_depth_map* map);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* R3DG_H_ */
