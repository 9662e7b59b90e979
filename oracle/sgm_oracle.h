/*
 * CPU oracle for the SGM stereo path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain C++/OpenMP restatement of the reference kernels' integer
 * arithmetic (ethan-li-coding/FasterStereoCuda, SgmStereoCuda/cusgm_*.cu).
 * It is the checker for the HIP path and the CPU baseline in the benchmark.
 * Nothing in the product library links or calls it.
 *
 * Every function names the reference lines it follows.  Where the reference
 * behaviour is racy or undefined, the oracle implements one documented
 * decision (ids A1..A17), and the HIP kernels implement the same decision.
 *
 * Layouts (all dense, row-major, no pitch):
 *   image    u8  [H][W]
 *   census   u64 [H][W]     (5x5 census uses the low 25 bits)
 *   cost     u8  [H][W][D]  (d fastest)
 *   path     u8  [H][W][D]  one volume per aggregation direction
 *   disp     f32 [H][W]     invalid = -999999.0f
 */
#ifndef SGM_ORACLE_H_
#define SGM_ORACLE_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_INVALID_DISP (-999999.0f)
#define ORC_INVALID_COST 127u

/* aggregation directions, in the order the reference stores them
 * (cusgm_ca.cu:4101-4108): 0 top->bottom, 1 left->right, 2 bottom->top,
 * 3 right->left */
enum { ORC_DIR_UD = 0, ORC_DIR_LR = 1, ORC_DIR_DU = 2, ORC_DIR_RL = 3 };

typedef struct orc_options {
    int32_t cs_type;          /* 0 = 5x5, 1 = 9x7 */
    int32_t p1;
    int32_t p2;
    float   uniqueness;       /* CuSGMOption::unique_threshold */
    int32_t do_lr_check;
    float   lr_threshold;
    int32_t do_median;
    int32_t do_remove_peaks;  /* left map only, after the filters */
    float   peaks_ratio;      /* CuSGMOption::peaks_ratio_threshold */
    int32_t peak_strips;      /* row strips (the reference: its host's core count) */
} orc_options;

int  orc_num_threads(void);

void orc_census(const uint8_t* img, int W, int H, int cs_type, int is_right,
                uint64_t* out);

void orc_cost(const uint64_t* cs_base, const uint64_t* cs_match, int W, int H,
              int min_d, int D, int left_to_right, int cs_type,
              const uint8_t* l2r_cost, uint8_t* cost);

void orc_aggregate(const uint8_t* cost, int W, int H, int D, int roi_x,
                   int dir, int p1, int p2, uint8_t* out);

void orc_wta(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2,
             const uint8_t* p3, int W, int H, int D, int min_d, int roi_x,
             float uniqueness, float* disp);

void orc_lrcheck(float* disp_l, const float* disp_r_in, float* disp_r_out,
                 int W, int H, float threshold);

void orc_median3x3(const float* in, float* out, int W, int H);

/* Peak removal (StereoCudaImpl.cpp:32-97, 463-498), in place: per strip of
 * H / strips rows (rows past the last strip untouched), 4-connected segments
 * of valid pixels joined when |d - d'| <= 1 (float) smaller than
 * (int)((float)(W * rows) * ratio) pixels become invalid.  Returns the
 * number of pixels removed. */
int orc_remove_peaks(float* disp, int W, int H, int strips, float ratio);

/* ---- coarse-to-fine pyramid (FasterStereoCuda) ---- */
#define ORC_INVALID_SHORT (-32768)

void orc_zoom_half(const uint8_t* src, int w, int h, uint8_t* dst);
void orc_init_next(const float* coarse, int w, int h, int level_range, int16_t* next);
void orc_cost_init(const uint64_t* cl, const uint64_t* cr, const int16_t* init, int W, int H,
                   int level_range, int cs_type, uint8_t* cost);
void orc_aggregate_init(const uint8_t* cost, const int16_t* init, int W, int H,
                        int level_range, int dir, int p1, int p2, uint8_t* out);
void orc_wta_init(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2, const uint8_t* p3,
                  const int16_t* init, int W, int H, int level_range, float uniqueness,
                  float* disp);
void orc_dilation(const float* in, float* out, int W, int H, int wnd);
int orc_dilation_window(int w, int h);
/* One finer level: census, init cost, 4 paths, init WTA, [median]. */
int orc_match_init(const uint8_t* left, const uint8_t* right, const int16_t* init, int W, int H,
                   int level_range, const orc_options* opt, float* disp);
/* Whole pyramid: num_layers levels, coarsest plain (with LR if opt->do_lr_check,
 * then a dilation when num_layers >= 2),
 * finer levels searched in a window of level_range around the propagated
 * disparity; uniqueness opt->uniqueness at level 0 and 1.0 above. */
int orc_pyramid(const uint8_t* left, const uint8_t* right, int W, int H, int min_disp,
                int max_disp, int num_layers, int level_range, const orc_options* opt,
                float* disp);

/* The same pyramid with an explicit coarsest interval [dmin, dmin + range). */
int orc_pyramid_at(const uint8_t* left, const uint8_t* right, int W, int H, int dmin, int range,
                   int num_layers, int level_range, const orc_options* opt, float* disp);
/* Coarsest interval from a disparity interval (FasterStereoCuda.cpp:265-281),
 * from a depth interval and camera (:373-391; cam = {x0_left, y0_left,
 * x0_right, y0_right, focus, baseline}), and the shared power-of-two rounding. */
void orc_disp_interval(int min_disp, int max_disp, int W, int num_layers, int* dmin, int* range);
int orc_depth_interval(double min_depth, double max_depth, int W, int num_layers,
                       const double cam[6], int* dmin, int* range);
void orc_round_interval(double span, int W, int* dmin, int* range);
/* Depth from disparity in float64 (cusgm_util.cu:7-33), invalid passed through. */
void orc_depth(const float* disp, size_t n, const double cam[6], float* depth);

/* Whole plain pipeline (StereoCudaImpl::Match order).  disp_r may be NULL.
 * Returns 0 on success. */
int orc_match(const uint8_t* left, const uint8_t* right, int W, int H,
              int min_d, int D, const orc_options* opt, float* disp_l,
              float* disp_r);

/* orc_match with a region of interest per view (x, y, w, h; NULL = the whole
 * matched region); see sgm_oracle.cpp. */
int orc_match_roi(const uint8_t* left, const uint8_t* right, int W, int H, int min_d, int D,
                  const orc_options* opt, const int* roi_l, const int* roi_r, float* disp_l,
                  float* disp_r);

#ifdef __cplusplus
}
#endif
#endif
