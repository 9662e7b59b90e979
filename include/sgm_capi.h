/*
 * sgm_capi.h -- C ABI of the MI355X SGM stereo engine (libfsgm.so).
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no C++ or torch
 * types.  The C++ drop-in classes (StereoCuda.h, FasterStereoCuda.h) and the
 * Python bindings call only these functions.
 *
 * Conventions
 *   - return value: SGM_OK (0) or a negative SGM_E_* code; the message of
 *     the last failure on a handle is in sgm_last_error(handle) (and
 *     sgm_last_error(NULL) for calls without a handle).
 *   - images are tightly packed row-major uint8 W x H; disparity maps are
 *     tightly packed float32 W x H with invalid = -999999.0f.
 *   - "device" entry points take device pointers and a hipStream_t passed as
 *     void* (NULL = the null stream) and return after enqueueing work.
 *   - a handle is not thread safe; it belongs to the device that was current
 *     when sgm_init ran.
 */
#ifndef SGM_CAPI_H_
#define SGM_CAPI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGM_OK               0
#define SGM_E_INVALID_ARG   (-1)   /* bad size, pointer or option value */
#define SGM_E_HIP           (-2)   /* a HIP runtime call failed */
#define SGM_E_NOT_INIT      (-3)   /* handle used before a successful init */
#define SGM_E_UNSUPPORTED   (-4)   /* option outside the built scope */
#define SGM_E_NO_DEVICE     (-5)   /* no HIP device */

#define SGM_INVALID_FLOAT   (-999999.0f)
#define SGM_INVALID_SHORT   (-32768)
#define SGM_LEVEL_RANGE     32

/* Mirrors CuSGMOption field by field (SgmStereoCuda/sgm_types.h:21-65);
 * enums are int32, bools are int32 (0/1). */
typedef struct sgm_option {
    int32_t p1;
    int32_t p2_init;
    int32_t p2_max;                 /* read by nobody, as in the reference */
    int32_t cs_type;                /* 0 = 5x5 census, 1 = 9x7 census */
    int32_t num_paths;              /* read by nobody; 4 paths are used */
    int32_t do_median_filter;
    int32_t do_lr_check;
    float   lr_threshold;
    float   unique_threshold;
    int32_t using_constant_p2;      /* only constant P2 is built */
    int32_t do_remove_peaks;        /* left-map peak removal (handles and pyramids;
                                       batches refuse it); see sgm_set_peak_strips */
    float   peaks_ratio_threshold;  /* in [0, 1] */
    int32_t post_filter_type;       /* must be 0 = PF_NONE */
    int32_t morphology_type;        /* must be 0 = MP_NONE */
} sgm_option;

/* StereoROI_T (sgm_types.h:89-92). */
typedef struct sgm_roi { int32_t x, y, w, h; } sgm_roi;

typedef struct sgm_handle sgm_handle;

/* Replaces `new StereoCudaImpl` (SgmStereoCuda/StereoCuda.cpp:6-9). */
sgm_handle* sgm_create(void);
/* Replaces ~StereoCuda (StereoCuda.cpp:11-16): releases and frees. */
void sgm_destroy(sgm_handle* h);

/* Replaces StereoCuda::Init (StereoCuda.h:31, StereoCudaImpl.cpp:128-261).
 * Arguments are validated before the device is touched.  disp_range must be
 * 64 * 2^n with n in 0..3, or 32 (LEVEL_RANGE) for a handle that matches only
 * around an initial disparity map; P1 and P2 must be >= 0 with
 * P1 + P2 <= 65000.  using_constant_p2 is read only with an initial map, as in
 * the reference (cusgm_ca.cu:4005-4008). */
int sgm_init(sgm_handle* h, int32_t width, int32_t height, int32_t min_disparity,
             int32_t disp_range, const sgm_option* opt, int32_t print_log);

/* Replaces StereoCuda::Match (StereoCuda.h:44, StereoCudaImpl.cpp:507-546):
 * host pointers, blocking.  disp_right may be NULL.
 *   - init_disp_left NULL: the plain pass (disp_range 64 * 2^n handles).
 *   - init_disp_left set: a disp_range 32 handle matches in a 32-disparity
 *     window around each pixel's initial disparity (window cost, window
 *     aggregation, window WTA, median; StereoCudaImpl.cpp:363-446,
 *     cusgm_ca.cu:4052-4077).  That needs do_lr_check = 0, using_constant_p2
 *     and min_disparity <= 0; disp_right, if given, is all INVALID.
 * roi_left / roi_right (either may be NULL) restrict that view's aggregation
 * and WTA to the rectangle clipped to the image: the paths start on its
 * edges, the WTA's border rules stay those of the image, pixels outside are
 * INVALID, and the LR check, median and peak removal run on the full maps
 * (peaks over the left rectangle's rows).  A negative width or height is
 * SGM_E_INVALID_ARG; a region with init_disp_left is SGM_E_UNSUPPORTED. */
int sgm_match(sgm_handle* h, const uint8_t* img_left, const uint8_t* img_right,
              float* disp_left, float* disp_right, const int16_t* init_disp_left,
              const sgm_roi* roi_left, const sgm_roi* roi_right);

/* The plain pass on device-resident buffers, enqueued on `stream`; no host
 * synchronisation.  The kernels read the images and write the maps in place
 * (no staging copies).  disp_right may be NULL. */
int sgm_match_device(sgm_handle* h, const uint8_t* d_left, const uint8_t* d_right,
                     float* d_disp_left, float* d_disp_right, void* stream);

/* Camera of a rectified pair, field for field CamParam_T (sgm_types.h:69-84):
 * principal points (pixels), focal length (pixels) and baseline. */
typedef struct sgm_camera {
    double x0_left, y0_left;
    double x0_right, y0_right;
    double focus;
    double baseline;
} sgm_camera;

/* Replaces cusgm_util::ComputeDepthCuda (cusgm_util.cu:7-45):
 *   depth = (float)(baseline * focus / ((double)disp + (x0_right - x0_left)))
 * in float64, with the invalid value (-999999.0f) passed through.  Device
 * pointers, packed W x H, enqueued on `stream`. */
int sgm_depth_device(const float* d_disp, int32_t width, int32_t height, const sgm_camera* cam,
                     float* d_depth, void* stream);

/* Replaces StereoCuda::Init2 (StereoCuda.h:57, StereoCudaImpl.cpp:263-278):
 * sgm_init, and the camera used by sgm_match2. */
int sgm_init2(sgm_handle* h, int32_t width, int32_t height, int32_t min_disparity,
              int32_t disp_range, const sgm_option* opt, const sgm_camera* cam, int32_t print_log);

/* Replaces StereoCuda::Match2 (StereoCuda.h:69, StereoCudaImpl.cpp:548-583):
 * the sgm_match pipeline, then the depth of the final left map (as
 * sgm_depth_device) copied to depth_left.  Needs a handle set up by sgm_init2;
 * init_disp_left and the ROIs follow the rules of sgm_match. */
int sgm_match2(sgm_handle* h, const uint8_t* img_left, const uint8_t* img_right, float* depth_left,
               const int16_t* init_disp_left, const sgm_roi* roi_left, const sgm_roi* roi_right);

/* Replaces StereoCuda::SetMinDisparity / GetDispartyRange (StereoCuda.h:71-72). */
int sgm_set_min_disparity(sgm_handle* h, int32_t min_disparity);
int32_t sgm_get_disparity_range(const sgm_handle* h);

/* Replaces StereoCuda::Release (StereoCuda.h:75).  Safe to call twice. */
int sgm_release(sgm_handle* h);

const char* sgm_last_error(const sgm_handle* h);

/* Replaces StereoCuda::MallocPageLockedPtr / FreePageLockedPtr
 * (StereoCuda.h:94-97). */
int sgm_host_alloc(void** ptr, size_t size);
int sgm_host_free(void* ptr);

/* Replaces StereoCuda::GetRoiFromDispMap (StereoCuda.h:84,
 * StereoCudaImpl.cpp:585-628): host-side scan.  A NULL pointer or a size
 * <= 0 leaves `roi` untouched; a map without valid pixels leaves x and w. */
void sgm_roi_from_disparity(const float* disp, int32_t width, int32_t height, sgm_roi* roi);

/* Where the aggregation takes its matching cost (no reference counterpart;
 * the reference always stores the cost volume, cusgm_cc.cu:118-172):
 *   SGM_COST_ON_THE_FLY  recompute it from the census words inside the
 *                        aggregation wherever the plan allows (D <= 256, not
 *                        the 5x5 right view): no cost volume is built or read
 *   SGM_COST_STORED      always build and read the u8 cost volume
 *   SGM_COST_AUTO        on the fly when the plan also has enough scanlines
 *                        to fill the device (batches), stored otherwise
 *                        (single pairs finish sooner that way)
 * Process-wide; read when a handle, batch or pyramid is created.  All give
 * byte-identical results. */
#define SGM_COST_AUTO 0
#define SGM_COST_STORED 1
#define SGM_COST_ON_THE_FLY 2
int sgm_set_cost_mode(int32_t mode);
int32_t sgm_get_cost_mode(void);

/* How the window aggregation of the pyramid's finer levels (and
 * sgm_stage_aggregate_init) holds the disparities of a scanline (no
 * reference counterpart; cusgm_ca.cu:1181-1339 is the arithmetic):
 *   SGM_WINDOW_U32    one per lane in 32-bit arithmetic (any P1, P2)
 *   SGM_WINDOW_PAIRS  two per lane as packed u16 (P1 <= 127, P2 <= 254)
 *   SGM_WINDOW_QUADS  four per lane as two packed u16 pairs (same range)
 *   SGM_WINDOW_AUTO   the packed form with the fewest instructions that
 *                     still gives the device a wave per SIMD, the 32-bit
 *                     form outside the packed range
 * A packed form asked for outside its range runs the 32-bit form.
 * Process-wide, read at every launch.  All give byte-identical results. */
#define SGM_WINDOW_AUTO 0
#define SGM_WINDOW_U32 1
#define SGM_WINDOW_PAIRS 2
#define SGM_WINDOW_QUADS 3
int sgm_set_window_form(int32_t form);
int32_t sgm_get_window_form(void);

/* How the stored-cost aggregation reaches the WTA (no reference counterpart;
 * the reference runs the four directions, then the WTA, cusgm_ca.cu:4080-4135):
 *   SGM_WTA_FUSED     left->right, right->left and top->bottom in one launch,
 *                     then bottom->top fused with the WTA: no bottom->top
 *                     volume is written or read
 *   SGM_WTA_SEPARATE  the four directions in one launch, then the WTA: the
 *                     short bottom->top walk runs beside the long horizontal
 *                     ones instead of after them
 *   SGM_WTA_AUTO      SEPARATE when the four-direction launch has at most
 *                     2.5 waves per SIMD (single pairs, where the scanline
 *                     walks set the time) and no matched image is more than
 *                     2.5 times as wide as it is high; FUSED otherwise
 * The on-the-fly kernels and the persistent batch schedules always fuse.
 * Process-wide, read when the launches are enqueued (a batch's HIP graph keeps
 * the form it was captured with).  All give byte-identical results. */
#define SGM_WTA_AUTO 0
#define SGM_WTA_FUSED 1
#define SGM_WTA_SEPARATE 2
int sgm_set_wta_form(int32_t form);
int32_t sgm_get_wta_form(void);

/* Row strips of the peak removal (do_remove_peaks; StereoCudaImpl.cpp:32-97,
 * 463-498).  The reference splits the left map into
 * hardware_concurrency() * 3 / 4 strips of H / strips rows (rows past the last
 * strip are left alone) and removes, per strip, the 4-connected segments of
 * valid pixels (neighbours joined when |d - d'| <= 1) smaller than
 * (int)((float)(W * rows_per_strip) * peaks_ratio_threshold) pixels.  The
 * strip count therefore depends on the reference's host; here it is this
 * process-wide value, default 1 (the whole image).  Set it to the reference
 * host's count to reproduce that host.  Read at every match. */
int sgm_set_peak_strips(int32_t strips);
int32_t sgm_get_peak_strips(void);

/* StereoCuda::get_invad_float / get_invad_short / get_level_range
 * (StereoCuda.h:87-91). */
float sgm_invalid_float(void);
int16_t sgm_invalid_short(void);
int16_t sgm_level_range(void);

/* ---- batched matching (no reference counterpart) ---------------------------
 * Many pairs, each with its own size and disparity range, matched by one
 * launch per stage.  Pointers in the descriptors are device pointers and are
 * bound at creation; the batch owns every intermediate buffer.  Per pair the
 * output is byte-identical to sgm_match on a handle initialised with the
 * same size, range and options. */
typedef struct sgm_pair_desc {
    int32_t width, height, min_disparity, disp_range;
    const uint8_t* d_left;
    const uint8_t* d_right;
    float* d_disp_left;
    float* d_disp_right;            /* may be NULL */
} sgm_pair_desc;

typedef struct sgm_batch sgm_batch;

/* Returns NULL on failure (message in sgm_last_error(NULL)). */
sgm_batch* sgm_batch_create(const sgm_pair_desc* pairs, int32_t n, const sgm_option* opt);
/* Enqueues the whole pipeline for every pair on `stream`. */
int sgm_batch_run(sgm_batch* b, void* stream);
/* With timing on, every run records HIP events between its stages. */
int sgm_batch_set_timing(sgm_batch* b, int32_t on);
/* Sums of the per-stage times (ms) of the runs recorded since the last call:
 * ms[0..4] = census, cost, aggregate (left->right, right->left, top->bottom),
 * aggregate bottom->top fused with the WTA, post.  With the separate WTA form
 * (sgm_set_wta_form) ms[2] covers all four directions and ms[3] the WTA.
 * Waits for those runs. */
int sgm_batch_stage_times(sgm_batch* b, float* ms, int32_t* runs);
size_t sgm_batch_device_bytes(const sgm_batch* b);
void sgm_batch_destroy(sgm_batch* b);

/* Scanline schedule of the two aggregation launches.
 *   SGM_SCHED_GRID           one workgroup per 4-scanline group (default)
 *   SGM_SCHED_JOB_MAJOR      persistent: `waves` resident workgroups pull the
 *                            groups from a device queue, pair by pair
 *   SGM_SCHED_LONGEST_FIRST  persistent, queue sorted by work (scanline length
 *                            x disparity range), longest first
 * waves = 0 picks occupancy x compute units.  Results are identical under
 * every schedule; only the order in which scanline groups run changes. */
#define SGM_SCHED_GRID 0
#define SGM_SCHED_JOB_MAJOR 1
#define SGM_SCHED_LONGEST_FIRST 2
int sgm_batch_set_schedule(sgm_batch* b, int32_t policy, int32_t waves);
int sgm_batch_get_schedule(const sgm_batch* b, int32_t* policy, int32_t* waves);
/* Times the candidate schedules on `stream` (`reps` runs each, after one
 * warm-up run), keeps the fastest and remembers it for every later batch of
 * the same signature (device, options, per-pair size / range) in this
 * process; a remembered signature is applied without timing unless
 * `force` != 0.  Writes the chosen policy / waves and its ms per run. */
int sgm_batch_autotune(sgm_batch* b, void* stream, int32_t reps, int32_t force, int32_t* policy,
                       int32_t* waves, float* ms_per_run);
/* With graphs on, sgm_batch_run launches one HIP graph holding the whole
 * pipeline (captured on a private stream at the first run after a change of
 * schedule) instead of the individual kernels.  Timing and graphs exclude
 * each other. */
int sgm_batch_use_graph(sgm_batch* b, int32_t on);

/* ---- differentiable FP32 SGM (no reference counterpart) ------------------
 * Float cost volume C[B][H][W][D], per-pixel penalties P1, P2 [B][H][W],
 * soft-argmin disparity with temperature T.  Four directions over the whole
 * image with the textbook boundary (no d-1 term at d = 0, no d+1 term at
 * d = D-1).  D must be 64, 128 or 256.  All buffers are device memory owned
 * by the caller; sizes in floats come from the two query functions.  The
 * forward stores S (needed by the backward) and, in d_ckpt, what the
 * backward needs of each direction: the winning term of every cell (2 bits)
 * and the argmin of the previous step per pixel (1 byte), plus two floats
 * per pixel; about a quarter of one volume, and no per-direction volume.
 * The transient workspace (d_work, shared by forward and backward) is one
 * direction volume plus two per-pixel planes. */
size_t sgm_dsgm_checkpoint_floats(int32_t B, int32_t H, int32_t W, int32_t D);
size_t sgm_dsgm_workspace_floats(int32_t B, int32_t H, int32_t W, int32_t D);
int sgm_dsgm_forward(const float* d_cost, const float* d_p1, const float* d_p2, int32_t B,
                     int32_t H, int32_t W, int32_t D, float temperature, float* d_S,
                     float* d_disp, float* d_ckpt, float* d_work, void* stream);
/* Gradients of a scalar loss given dLoss/d disp.  dLoss/dS is formed from S
 * where it is used; d_grad_cost, d_grad_p1 and d_grad_p2 are built in place
 * over the directions.  The recurrences are reversed from what the forward
 * saved in d_ckpt: d_cost, d_p1 and d_p2 are checked but not read. */
int sgm_dsgm_backward(const float* d_cost, const float* d_p1, const float* d_p2,
                      const float* d_S, const float* d_disp, const float* d_ckpt,
                      const float* d_grad_disp, int32_t B, int32_t H, int32_t W, int32_t D,
                      float temperature, float* d_grad_cost, float* d_grad_p1,
                      float* d_grad_p2, float* d_work, void* stream);

/* ---- coarse-to-fine pyramid ---------------------------------------------
 * Replaces HierSgmCudaImpl (FasterStereoCuda/FasterStereoCuda.cpp:203-295
 * Init, :408-529 Match) with every level on the GPU.  Level i is
 * (W >> i) x (H >> i); the coarsest level runs the plain pipeline over
 * [min/2^i, max/2^i) rounded up to 2^k >= 64 (LR check there if
 * opt->do_lr_check); each finer level searches level_range (16 or 32)
 * disparities around 2 d_coarse - level_range/2.  opt->unique_threshold
 * applies at level 0, 1.0 above.  With opt->do_remove_peaks the coarsest
 * level's map loses its peaks after its filters (and after its dilation when
 * there are finer levels), with opt->peaks_ratio_threshold as given; the
 * reference's schedule (FasterStereoCuda.cpp:182-196) removes them only there,
 * with ratio 0.0005 / 0.001 / 0.005 for 1 / 2 / more levels.  NULL on
 * failure (sgm_last_error(NULL)). */
typedef struct sgm_pyramid sgm_pyramid;
sgm_pyramid* sgm_pyramid_create(int32_t width, int32_t height, int32_t min_disp, int32_t max_disp,
                                int32_t num_layers, int32_t level_range, const sgm_option* opt);
/* host pointers, blocking */
int sgm_pyramid_match(sgm_pyramid* p, const uint8_t* left, const uint8_t* right, float* disp);
/* device pointers, enqueued on stream; the finest level reads d_left /
 * d_right and writes d_disp in place (nothing is staged), so they must stay
 * valid until the stream reaches the end of the call's work */
int sgm_pyramid_match_device(sgm_pyramid* p, const uint8_t* d_left, const uint8_t* d_right,
                             float* d_disp, void* stream);
/* Replaces HierSgmCudaImpl::Init2 (FasterStereoCuda.cpp:297-406): as
 * sgm_pyramid_create, with the coarsest level's interval taken from a depth
 * interval.  With s = 2^(num_layers-1) and the camera scaled by 1/s
 * (principal points and focus; CamParam_T::Scale, sgm_types.h:75-82):
 *   min = baseline * focus / max_depth - x0_right + x0_left   (float64)
 *   max = baseline * focus / min_depth - x0_right + x0_left
 * the range is the smallest 64 * 2^k >= max - min (halved once if above
 * twice the width), min moves to width - range (not below -width) when it
 * reaches the width, and the level starts at floor(min).  The camera is kept
 * for sgm_pyramid_match2.  0 < min_depth < max_depth, all finite. */
sgm_pyramid* sgm_pyramid_create2(int32_t width, int32_t height, float min_depth, float max_depth,
                                 int32_t num_layers, int32_t level_range, const sgm_option* opt,
                                 const sgm_camera* cam);
/* Replaces HierSgmCudaImpl::Match2 (FasterStereoCuda.cpp:531-601): the
 * pyramid match, then the depth of the level-0 map with the unscaled camera.
 * Needs a pyramid from sgm_pyramid_create2. */
int sgm_pyramid_match2(sgm_pyramid* p, const uint8_t* left, const uint8_t* right, float* depth);
int sgm_pyramid_match2_device(sgm_pyramid* p, const uint8_t* d_left, const uint8_t* d_right,
                              float* d_depth, void* stream);
/* range of the coarsest level (and its minimum disparity through min_disp) */
int32_t sgm_pyramid_coarsest_range(const sgm_pyramid* p, int32_t* min_disp);
size_t sgm_pyramid_device_bytes(const sgm_pyramid* p);
void sgm_pyramid_destroy(sgm_pyramid* p);

/* ---- stage entry points (device pointers; used by the parity tests) ----
 * Each runs one stage of the plain pipeline over one view.  Layouts:
 * census u64 [H][W]; cost and path volumes u8 [H][W][D] (d fastest). */

/* cusgm_cc.cu:30-106. */
int sgm_stage_census(const uint8_t* d_img, int32_t width, int32_t height, int32_t cs_type,
                     int32_t is_right, uint64_t* d_census, void* stream);
/* cusgm_cc.cu:118-172, 238-293.  For the 5x5 right-to-left cost the pixels
 * the reference kernel skips (cusgm_cc.cu:288) get the left-to-right cost
 * its reused buffer held, recomputed from the two census maps. */
int sgm_stage_cost(const uint64_t* d_census_base, const uint64_t* d_census_match,
                   int32_t width, int32_t height, int32_t min_disparity, int32_t disp_range,
                   int32_t left_to_right, int32_t cs_type, uint8_t* d_cost, void* stream);
/* cusgm_ca.cu:1480-1888 and the generic kernels; d_paths[r] for r = 0 top->bottom,
 * 1 left->right, 2 bottom->top, 3 right->left. */
int sgm_stage_aggregate(const uint8_t* d_cost, int32_t width, int32_t height,
                        int32_t min_disparity, int32_t disp_range, int32_t p1, int32_t p2,
                        uint8_t* const d_paths[4], void* stream);
/* The aggregation the pipeline runs when the cost comes from the census
 * words on the fly (sgm_set_cost_mode): cusgm_cc.cu:118-172 cost inside
 * cusgm_ca.cu:1480-1888 recurrence, then the WTA (cusgm_ca.cu:3606-3809)
 * fused with the bottom->top direction.  One view: the base census is this
 * view's, is_right = 0 matches x - d (left view), 1 matches x + d.  Writes
 * d_paths[0] (top->bottom), d_paths[1] (left->right), d_paths[3]
 * (right->left) and the raw disparity map; d_paths[2] is not written (the
 * fused pass keeps bottom->top in registers) and may be NULL.  D <= 256,
 * not the 5x5 right view (SGM_E_UNSUPPORTED). */
int sgm_stage_aggregate_wta_otf(const uint64_t* d_census_base, const uint64_t* d_census_match,
                                int32_t width, int32_t height, int32_t min_disparity,
                                int32_t disp_range, int32_t is_right, int32_t cs_type, int32_t p1,
                                int32_t p2, float uniqueness, uint8_t* const d_paths[4],
                                float* d_disp, void* stream);
/* cusgm_ca.cu:3149-3232, 3606-3809. */
int sgm_stage_wta(uint8_t* const d_paths[4], int32_t width, int32_t height,
                  int32_t min_disparity, int32_t disp_range, float uniqueness,
                  float* d_disp, void* stream);
/* cusgm_ca.cu:3873-3898 and cusgm_df.cu:16-75: LR check (if do_lr_check)
 * then 3x3 median (if do_median); d_out_right (may be NULL) receives the
 * right map with the LR invalidations applied. */
int sgm_stage_post(const float* d_raw_left, const float* d_raw_right, int32_t width,
                   int32_t height, int32_t do_lr_check, float lr_threshold,
                   int32_t do_median, float* d_out_left, float* d_out_right, void* stream);

/* StereoCudaImpl.cpp:32-97, 463-498: peak removal of a device map in place
 * (as with do_remove_peaks; strips as in sgm_set_peak_strips).  d_scratch
 * holds 2 * width * height ints. */
int sgm_stage_remove_peaks(float* d_disp, int32_t width, int32_t height, int32_t strips,
                           float ratio, int32_t* d_scratch, void* stream);

/* No counterpart in the reference.  sgm_stage_remove_peaks without the
 * workspace, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * union-find scratch is a fixed array inside the library, one per device,
 * for maps of at most 2^21 = 2097152 pixels (16 MiB, held until the process
 * ends); a larger map returns SGM_E_INVALID_ARG.  No allocation and no host
 * synchronisation: the work is enqueued on `stream`.  Calls on one device
 * share that scratch, so calls that overlap in time on one device race. */
int sgm_stage_remove_peaks_noalloc(float* d_disp, int32_t width, int32_t height, int32_t strips,
                                   float ratio, void* stream);

/* FasterStereoCuda.cpp:114-136: d_dst[(h/2)][(w/2)] = d_src[2i][2j]. */
int sgm_stage_zoom(const uint8_t* d_src, int32_t w, int32_t h, uint8_t* d_dst, void* stream);
/* FasterStereoCuda.cpp:138-166: base disparities of a w x h level from the
 * (w/2) x (h/2) coarse map. */
int sgm_stage_init_next(const float* d_coarse, int32_t w, int32_t h, int32_t level_range,
                        int16_t* d_next, void* stream);
/* cusgm_cc.cu:173-237 / 294-360: window cost [h][w][level_range]. */
int sgm_stage_cost_init(const uint64_t* d_census_l, const uint64_t* d_census_r,
                        const int16_t* d_init, int32_t w, int32_t h, int32_t level_range,
                        int32_t cs_type, uint8_t* d_cost, void* stream);
/* cusgm_ca.cu:1181-1339: window aggregation of the 4 directions. */
int sgm_stage_aggregate_init(const uint8_t* d_cost, const int16_t* d_init, int32_t w, int32_t h,
                             int32_t level_range, int32_t p1, int32_t p2,
                             uint8_t* const d_paths[4], void* stream);
/* cusgm_ca.cu:3425-3512: window WTA. */
int sgm_stage_wta_init(uint8_t* const d_paths[4], const int16_t* d_init, int32_t w, int32_t h,
                       int32_t level_range, float uniqueness, float* d_disp, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SGM_CAPI_H_ */
