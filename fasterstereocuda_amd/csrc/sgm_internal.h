// Internal types shared by the kernels and the C-ABI layer.
//
// Work is described as a list of "jobs": one job is one view of one stereo
// pair (the left view matched against the right, or the right view against
// the left).  Every stage kernel walks a flat item space; per stage the host
// builds a prefix-sum table offs[0..njobs] of items per job, and a workgroup
// (or wave) finds its job by binary search.  The same kernels therefore run
// one pair (the StereoCuda::Match path), a uniform batch, or a batch of
// pairs with different W, H and D.
#pragma once
#include <stdint.h>

#include <vector>

#include <hip/hip_runtime.h>

#include "sgm_capi.h"

namespace fsgm {

constexpr float kInvalidDisp = -999999.0f;
constexpr int kInvalidCost = 127;
constexpr int kMaxD = 512;

enum Dir { kUD = 0, kLR = 1, kDU = 2, kRL = 3 };

struct SgmJob {
    const uint8_t* img;            // this view, u8 [H][W]
    uint64_t* census;              // this view's census, u64 [H][W]
    const uint64_t* census_other;  // the other view's census
    uint8_t* cost;                 // u8 [H][W][D]
    uint8_t* path[4];              // u8 [H][W][D] per direction
    float* disp;                   // WTA output, f32 [H][W]
    int32_t W, H, D, min_d;
    int32_t roi_x;                 // max(0, min_d)
    int32_t is_right;              // 0: left view (matches x - d), 1: right view (x + d)
    int32_t cs_type;               // 0 = 5x5, 1 = 9x7
    int32_t p1, p2;
    float uniqueness;
    // WTA frame: the job is the block of an fW x fH image whose (0, 0) is
    // image pixel (fx0, fy0) -- a region of interest.  fW = 0: the job is
    // the whole image.  Border rules then use image coordinates and the map
    // written is the image's (row pitch fW).
    int32_t fx0, fy0, fW, fH;
};

struct SgmPost {                   // one stereo pair: LR check + median
    const float* raw_l;
    const float* raw_r;            // may be null when the LR check is off
    float* out_l;
    float* out_r;                  // may be null
    int32_t W, H;
    int32_t roi_x;                 // columns below are INVALID
    int32_t do_lr, do_median;
    float lr_thr;
};

// Items per job for each stage.
inline int census_items(const SgmJob& j) { return ((j.W + 63) / 64) * ((j.H + 15) / 16); }
constexpr int kCostSeg = 256;   // pixels of one row per cost workgroup
inline int cost_items(const SgmJob& j) { return j.H * ((j.W - j.roi_x + kCostSeg - 1) / kCostSeg); }
inline int agg_items(const SgmJob& j) { return 2 * ((j.H + 3) / 4) + 2 * ((j.W - j.roi_x + 3) / 4); }
inline int agg3_items(const SgmJob& j) { return 2 * ((j.H + 3) / 4) + ((j.W - j.roi_x + 3) / 4); }
inline int dw_items(const SgmJob& j) { return (j.W - j.roi_x + 3) / 4; }
inline int wta_items(const SgmJob& j) { return j.H * ((j.W + 15) / 16); }
inline int post_items(const SgmPost& p) { return ((p.W + 63) / 64) * ((p.H + 15) / 16); }
// On-the-fly cost aggregation (sgm_wide.hip).  Each kernel holds DPL
// disparities per lane, so a scanline takes D / DPL lanes (at most one
// 16-lane DPP row) and a wave covers 64 * DPL / D scanlines.
enum WideKind { kWideH = 0, kWideV = 1, kWideF = 2 };   // LR + RL, UD, DU + WTA
#ifndef FSGM_DPL_H
#define FSGM_DPL_H 16
#endif
#ifndef FSGM_DPL_V
#define FSGM_DPL_V 16
#endif
#ifndef FSGM_DPL_F
#define FSGM_DPL_F 16
#endif
__host__ __device__ constexpr int wide_dpl(int kind, int D) {
    const int want = kind == kWideH ? FSGM_DPL_H : kind == kWideV ? FSGM_DPL_V : FSGM_DPL_F;
    int dpl = want * 16 >= D ? want : D / 16;   // D / DPL <= 16
    // horizontal: a block of DPL steps spreads its census loads over the
    // line's D / DPL lanes, so DPL >= D / DPL
    if (kind == kWideH)
        while (dpl * dpl < D) dpl *= 2;
    return dpl;
}
__host__ __device__ constexpr int wide_lines(int kind, int D) { return 64 * wide_dpl(kind, D) / D; }
inline int wide_h_items(const SgmJob& j) {
    const int l = wide_lines(kWideH, j.D);
    return 2 * ((j.H + l - 1) / l);
}
inline int wide_v_items(const SgmJob& j, int kind) {
    const int l = wide_lines(kind, j.D);
    return (j.W - j.roi_x + l - 1) / l;
}
// Whether a job's cost can come from the census on the fly: D <= 256, and not
// the 5x5 right view, whose skipped-pixel rule needs the stored cost.
inline bool wide_ok(const SgmJob& j) { return j.D <= 256 && !(j.is_right && j.cs_type == 0); }

// Caller buffers bound for one launch sequence in place of the plan's own:
// the images of jobs 0 and 1 (the left and right view of a single-pair
// plan) and the outputs of post 0.  With `on` = 0 the plan's buffers are
// used; out_r may be null (no right output).  Passed to the census and post
// kernels by value, so binding other buffers needs no plan upload.
struct IoBind {
    const uint8_t* img[2];
    float* out_l;
    float* out_r;
    int32_t on;
};

// Device-side launch plan: jobs, per-stage prefix sums and totals.
struct StagePlan {
    SgmJob* jobs = nullptr;       // device
    int* offs_census = nullptr;   // device, njobs + 1 each
    int* offs_cost = nullptr;
    int* offs_agg = nullptr;
    int* offs_wta = nullptr;
    int* offs_agg3 = nullptr;     // three directions
    int* offs_dw = nullptr;       // bottom->top fused with the WTA
    int* offs_wh = nullptr;       // on-the-fly cost: left->right + right->left groups
    int* offs_wv = nullptr;       // on-the-fly cost: top->bottom groups
    int* offs_wf = nullptr;       // on-the-fly cost: bottom->top + WTA groups
    SgmPost* posts = nullptr;     // device
    int* offs_post = nullptr;
    int njobs = 0, nposts = 0;
    int njobs_ctw = 0;            // if > 0: jobs [0, njobs_ctw) run cost..wta
    int n_census = 0, n_cost = 0, n_agg = 0, n_wta = 0, n_post = 0, n_agg3 = 0, n_dw = 0;
    int n_wh = 0, n_wv = 0, n_wf = 0;
    int dmax = 64;
    bool wide = false;            // matching uses the on-the-fly cost kernels
    bool squarish = false;        // every matched job has W - roi_x <= 2.5 H
    IoBind io{};                  // host-side: bound caller buffers, if any
};

// Scanline schedule of the persistent aggregation launches: packed items
// (job << 16 | direction << 14 | group) in execution order, two counters.
struct AggSchedule {
    const uint32_t* items3 = nullptr;   // LR, RL, UD items
    int n3 = 0;
    int n3h = 0;                        // on-the-fly plans: items3 starts with n3h LR / RL items
    const uint32_t* items_dw = nullptr; // bottom->top + WTA items
    int ndw = 0;
    uint32_t* counters = nullptr;       // [3], reset by the first launch
    int grid = 0;                       // resident waves
};

// Launchers (sgm_kernels.hip).  All enqueue on `stream`.
hipError_t launch_census(const StagePlan& p, hipStream_t s);
hipError_t launch_cost(const StagePlan& p, hipStream_t s);
hipError_t launch_aggregate(const StagePlan& p, hipStream_t s);        // 4 directions
hipError_t launch_aggregate3(const StagePlan& p, hipStream_t s);       // LR, RL, UD
hipError_t launch_aggregate_wta(const StagePlan& p, hipStream_t s);    // DU + WTA
hipError_t launch_aggregate_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s);
hipError_t launch_aggregate_wta_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s);
// Resident 64-lane workgroups per compute unit of the persistent kernels.
hipError_t persistent_occupancy(const StagePlan& p, int* per_cu);
hipError_t launch_reset_counters(uint32_t* counters, int n, hipStream_t s);
// sgm_wide.hip: aggregation with the cost computed from the census words
hipError_t launch_wide3(const StagePlan& p, hipStream_t s);
hipError_t launch_wide_wta(const StagePlan& p, hipStream_t s);
hipError_t launch_wide3_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s);
hipError_t launch_wide_wta_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s);
hipError_t wide_occupancy(const StagePlan& p, int* per_cu);
// Process-wide SGM_COST_* mode, read when a plan is built; choose_wide
// applies it to a plan whose views all allow the on-the-fly cost
// (`eligible`) and that has n_wh horizontal scanline groups.
int cost_mode();
bool choose_wide(bool eligible, int n_wh);
hipError_t launch_wta(const StagePlan& p, hipStream_t s);
hipError_t launch_post(const StagePlan& p, hipStream_t s);
hipError_t launch_right_invalidate(const StagePlan& p, hipStream_t s);
// peaks.hip: segments of fewer than (int)((float)(W r) ratio) pixels of each
// r = H / strips row strip become invalid (in place); scratch = 2 W H ints
// Largest map (pixels) the allocation-free peak removal takes, and the
// current device's address of its fixed scratch (looked up once per device).
constexpr int kPeaksNoAllocMaxPixels = 1 << 21;
hipError_t peaks_noalloc_scratch(int** out);
hipError_t launch_remove_peaks(float* disp, int* scratch, int W, int H, int strips, float ratio,
                               hipStream_t s);
int peaks_threshold(int W, int strip_rows, float ratio);
int peak_strips();   // process-wide strip count (sgm_set_peak_strips)
// depth = (float)(f1 / ((double)disp + f2)), invalid kept (n pixels)
// rows [y0, y0 + hc) x columns [x0, x0 + wc) of a [H][W][D] u8 volume copied
// into a packed [hc][wc][D] volume (D a multiple of 16)
hipError_t launch_crop_volume(const uint8_t* src, uint8_t* dst, int W, int D, int x0, int y0,
                              int wc, int hc, hipStream_t s);
// columns [0, ncols) of a W x H float map set to INVALID
hipError_t launch_invalid_columns(float* disp, int W, int H, int ncols, hipStream_t s);
hipError_t launch_depth(const float* disp, float* depth, size_t n, double f1, double f2,
                        hipStream_t s);

// sgm_capi.hip: the matching sequence shared by the plain, batched and
// pyramid paths (marks: 6 events or null)
void restrict_matching(StagePlan* P, const std::vector<SgmJob>& jobs, int n);
hipError_t launch_match(const StagePlan& plan, bool right_invalidation, hipStream_t s,
                        hipEvent_t* marks, const AggSchedule* sched = nullptr);

}  // namespace fsgm
