// Coarse-to-fine pyramid (FasterStereoCuda) on the GPU, end to end.
//
// Mirrors HierSgmCudaImpl (ethan-li-coding/FasterStereoCuda,
// FasterStereoCuda/FasterStereoCuda.cpp:203-295 Init, :408-529 Match):
//   level i has size (W >> i) x (H >> i), made by 2x decimation;
//   the coarsest level runs the plain SGM over [min/2^i, max/2^i) rounded up
//   to a power of two >= 64, with the LR check if requested;
//   every finer level searches level_range disparities around
//   b = trunc(2 d_coarse - level_range / 2) (window cost + CostAdjust,
//   window aggregation, window WTA), then the 3x3 median;
//   uniqueness is the option's value at level 0 and 1.0 above (:176).
// The reference decimates and propagates on the CPU with a host round trip
// per level; here every step is a kernel on one stream.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pyr_internal.h"
#include "sgm_capi.h"

using namespace fsgm;

namespace {

inline size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }

#define PYR_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return set_global_error(SGM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

struct sgm_pyramid {
    int device = -1;
    sgm_option opt{};
    int W = 0, H = 0, layers = 1, LR = 32;
    int top_min = 0, top_range = 64;
    std::vector<int> lw, lh;
    void* arena = nullptr;
    size_t arena_size = 0;
    hipStream_t stream = nullptr;
    std::vector<uint8_t*> il, ir;
    std::vector<uint64_t*> cl, cr;
    std::vector<float*> out, raw;
    std::vector<int16_t*> init;
    std::vector<uint8_t*> cost;
    std::vector<std::array<uint8_t*, 4>> path;
    StagePlan top_plan;
    void* top_block = nullptr;
    std::vector<StagePlan> lvl_plan;
    std::vector<void*> lvl_block;
    bool has_cam = false;         // set by sgm_pyramid_create2 (level-0 camera)
    sgm_camera cam{};
    int* peaks = nullptr;         // peak-removal scratch of the coarsest level
};

namespace {

void free_pyramid(sgm_pyramid* p) {
    if (p->device >= 0) (void)hipSetDevice(p->device);
    if (p->stream) { (void)hipStreamSynchronize(p->stream); (void)hipStreamDestroy(p->stream); }
    if (p->arena) (void)hipFree(p->arena);
    if (p->top_block) (void)hipFree(p->top_block);
    if (p->peaks) (void)hipFree(p->peaks);
    for (void* b : p->lvl_block)
        if (b) (void)hipFree(b);
    delete p;
}

SgmJob census_job(const uint8_t* img, uint64_t* cs, const uint64_t* other, int w, int h,
                  int cs_type, int is_right) {
    SgmJob j{};
    j.img = img; j.census = cs; j.census_other = other;
    j.W = w; j.H = h; j.D = 64; j.min_d = 0; j.roi_x = 0;
    j.is_right = is_right; j.cs_type = cs_type;
    return j;
}

}  // namespace

namespace {

// Checks shared by both constructors; returns false with the error set.
bool pyramid_args_ok(int32_t width, int32_t height, int32_t num_layers, int32_t level_range,
                     const sgm_option* opt) {
    if (width <= 0 || height <= 0) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: bad size");
        return false;
    }
    if (level_range != 16 && level_range != 32) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: level_range must be 16 or 32");
        return false;
    }
    const int layers = std::max(1, num_layers);
    if ((height >> (layers - 1)) < 8 || (width >> (layers - 1)) < 8) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: too many layers for the image size");
        return false;
    }
    if (!opt || opt->cs_type < 0 || opt->cs_type > 1 || opt->p1 < 0 || opt->p2_init < 0 ||
        (int64_t)opt->p1 + opt->p2_init > 65000) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: bad options");
        return false;
    }
    if (opt->post_filter_type != 0 || opt->morphology_type != 0) {
        set_global_error(SGM_E_UNSUPPORTED, "pyramid: post filters and morphology are not built");
        return false;
    }
    // the finer levels aggregate around initial disparities, which the
    // reference does only with constant P2 (cusgm_ca.cu:4052-4077); a single
    // level is the plain path, which ignores the flag
    if (layers > 1 && !opt->using_constant_p2) {
        set_global_error(SGM_E_UNSUPPORTED,
                         "pyramid: levels above one need using_constant_p2 (the reference's "
                         "window aggregation has no adaptive-P2 form)");
        return false;
    }
    if (opt->do_remove_peaks && (int64_t)width * height > INT32_MAX) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: peak removal needs width * height < 2^31");
        return false;
    }
    if (opt->do_remove_peaks &&
        !(opt->peaks_ratio_threshold >= 0.0f && opt->peaks_ratio_threshold <= 1.0f)) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: peaks_ratio_threshold must be in [0, 1]");
        return false;
    }
    return true;
}

// Coarsest-level interval, shared by both constructors
// (FasterStereoCuda.cpp:266-281 and :373-391): the smallest 64 * 2^k covering
// span, halved once if above twice the (full) width; a minimum at or past the
// width moves to width - range, not below -width.
void coarsest_interval(double span, int width, int* dmin, int* range) {
    int r = 64;
    while (r < span && r < (1 << 20)) r <<= 1;   // beyond 512 the caller rejects it
    if (r > 2 * width) r >>= 1;
    *range = r;
    if (*dmin >= width) *dmin = std::max(width - r, -width);
}

sgm_pyramid* build_pyramid(int32_t width, int32_t height, int32_t num_layers, int32_t level_range,
                           const sgm_option* opt, int dmin, int range) {
    const int layers = std::max(1, num_layers);
    sgm_pyramid* p = new (std::nothrow) sgm_pyramid();
    if (!p) return nullptr;
    p->opt = *opt;
    p->W = width; p->H = height; p->layers = layers; p->LR = level_range;
    if (hipGetDevice(&p->device) != hipSuccess) {
        set_global_error(SGM_E_NO_DEVICE, "no HIP device");
        delete p;
        return nullptr;
    }
    const int top = layers - 1;
    for (int i = 0; i < layers; ++i) { p->lw.push_back(width >> i); p->lh.push_back(height >> i); }
    if (range > 512 || std::max(0, dmin) >= p->lw[top]) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: coarsest disparity range not supported");
        delete p;
        return nullptr;
    }
    p->top_min = dmin; p->top_range = range;
    // memory
    const bool lr = opt->do_lr_check != 0;
    size_t total = 0;
    for (int i = 0; i < layers; ++i) {
        const size_t npix = (size_t)p->lw[i] * p->lh[i];
        total += 2 * a256(npix) + 2 * a256(npix * 8) + 2 * a256(npix * 4);
        if (i == top) total += (lr ? 2 : 1) * 5 * a256(npix * range) + a256(npix * 4);
        else total += a256(npix * 2) + 5 * a256(npix * level_range);
    }
    // peaks are removed at the coarsest level only (FasterStereoCuda.cpp:182-196)
    const size_t top_pix = (size_t)p->lw[top] * p->lh[top];
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&p->arena, total) != hipSuccess ||
        (opt->do_remove_peaks && hipMalloc(&p->peaks, 2 * top_pix * sizeof(int)) != hipSuccess)) {
        set_global_error(SGM_E_HIP, "pyramid: device allocation failed");
        free_pyramid(p);
        return nullptr;
    }
    p->arena_size = total;
    uint8_t* cur = static_cast<uint8_t*>(p->arena);
    auto take = [&](size_t n) { uint8_t* r = cur; cur += a256(n); return r; };
    p->il.resize(layers); p->ir.resize(layers); p->cl.resize(layers); p->cr.resize(layers);
    p->out.resize(layers); p->raw.resize(layers); p->init.resize(layers, nullptr);
    p->cost.resize(layers, nullptr); p->path.resize(layers);
    p->lvl_plan.resize(layers); p->lvl_block.resize(layers, nullptr);
    for (int i = 0; i < layers; ++i) {
        const size_t npix = (size_t)p->lw[i] * p->lh[i];
        p->il[i] = take(npix); p->ir[i] = take(npix);
        p->cl[i] = reinterpret_cast<uint64_t*>(take(npix * 8));
        p->cr[i] = reinterpret_cast<uint64_t*>(take(npix * 8));
        p->out[i] = reinterpret_cast<float*>(take(npix * 4));
        p->raw[i] = reinterpret_cast<float*>(take(npix * 4));
    }
    // coarsest level: the plain job-table pipeline
    {
        const int w = p->lw[top], h = p->lh[top];
        const size_t npix = (size_t)w * h, ncell = npix * range;
        std::vector<SgmJob> jobs;
        SgmPost post{};
        for (int v = 0; v < 2; ++v) {
            SgmJob j = census_job(v ? p->ir[top] : p->il[top], v ? p->cr[top] : p->cl[top],
                                  v ? p->cl[top] : p->cr[top], w, h, opt->cs_type, v);
            j.D = range; j.min_d = dmin; j.roi_x = std::max(0, dmin);
            j.p1 = opt->p1; j.p2 = opt->p2_init;
            j.uniqueness = (top > 0) ? 1.0f : opt->unique_threshold;
            if (v == 0 || lr) {
                j.cost = take(ncell);
                for (int r = 0; r < 4; ++r) j.path[r] = take(ncell);
                j.disp = v ? reinterpret_cast<float*>(take(npix * 4)) : p->raw[top];
            }
            (v ? post.raw_r : post.raw_l) = j.disp;
            jobs.push_back(j);
        }
        post.out_l = p->out[top];
        post.W = w; post.H = h;
        post.roi_x = std::max(0, dmin);
        post.do_lr = lr ? 1 : 0;
        post.do_median = opt->do_median_filter ? 1 : 0;
        post.lr_thr = opt->lr_threshold;
        std::vector<SgmPost> posts{post};
        std::string err;
        if (upload_plan(jobs, posts, &p->top_plan, &p->top_block, &err)) {
            set_global_error(SGM_E_HIP, err);
            free_pyramid(p);
            return nullptr;
        }
        if (!lr) restrict_matching(&p->top_plan, jobs, 1);
    }
    // finer levels: census jobs + median post per level, window buffers
    for (int i = 0; i < top; ++i) {
        const int w = p->lw[i], h = p->lh[i];
        const size_t npix = (size_t)w * h, ncell = npix * level_range;
        p->init[i] = reinterpret_cast<int16_t*>(take(npix * 2));
        p->cost[i] = take(ncell);
        for (int r = 0; r < 4; ++r) p->path[i][r] = take(ncell);
        std::vector<SgmJob> jobs{census_job(p->il[i], p->cl[i], p->cr[i], w, h, opt->cs_type, 0),
                                 census_job(p->ir[i], p->cr[i], p->cl[i], w, h, opt->cs_type, 1)};
        SgmPost post{};
        post.raw_l = p->raw[i]; post.out_l = p->out[i];
        post.W = w; post.H = h; post.do_lr = 0;
        post.do_median = opt->do_median_filter ? 1 : 0;
        std::vector<SgmPost> posts{post};
        std::string err;
        if (upload_plan(jobs, posts, &p->lvl_plan[i], &p->lvl_block[i], &err)) {
            set_global_error(SGM_E_HIP, err);
            free_pyramid(p);
            return nullptr;
        }
    }
    return p;
}

}  // namespace

extern "C" {

sgm_pyramid* sgm_pyramid_create(int32_t width, int32_t height, int32_t min_disp, int32_t max_disp,
                                int32_t num_layers, int32_t level_range, const sgm_option* opt) {
    if (max_disp <= min_disp) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: bad disparity interval");
        return nullptr;
    }
    if (!pyramid_args_ok(width, height, num_layers, level_range, opt)) return nullptr;
    // FasterStereoCuda.cpp:265-281: integer division by the coarsest scale
    const int scale = 1 << (std::max(1, num_layers) - 1);
    int dmin = min_disp / scale, range = 0;
    coarsest_interval((double)(max_disp / scale - dmin), width, &dmin, &range);
    return build_pyramid(width, height, num_layers, level_range, opt, dmin, range);
}

sgm_pyramid* sgm_pyramid_create2(int32_t width, int32_t height, float min_depth, float max_depth,
                                 int32_t num_layers, int32_t level_range, const sgm_option* opt,
                                 const sgm_camera* cam) {
    if (!cam || !(min_depth > 0.0f) || !(max_depth > min_depth) || !std::isfinite(max_depth) ||
        !std::isfinite(cam->focus) || !std::isfinite(cam->baseline) ||
        !std::isfinite(cam->x0_left) || !std::isfinite(cam->x0_right)) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: bad camera or depth interval");
        return nullptr;
    }
    if (!pyramid_args_ok(width, height, num_layers, level_range, opt)) return nullptr;
    // FasterStereoCuda.cpp:373-391 with the camera scaled by 1/2^top
    // (CamParam_T::Scale, sgm_types.h:75-82), all in float64
    const double s = (double)(1 << (std::max(1, num_layers) - 1));
    const double x0l = cam->x0_left / s, x0r = cam->x0_right / s, f = cam->focus / s;
    const double dmin_f = cam->baseline * f / max_depth - x0r + x0l;
    const double dmax_f = cam->baseline * f / min_depth - x0r + x0l;
    if (!std::isfinite(dmin_f) || !std::isfinite(dmax_f) || dmin_f < -1e9 || dmax_f > 1e9) {
        set_global_error(SGM_E_INVALID_ARG, "pyramid: depth interval maps to no disparity interval");
        return nullptr;
    }
    // the reference tests the float64 minimum against the width and floors
    // it afterwards; floor(x) >= width exactly when x >= width
    int dmin = (int)std::floor(dmin_f), range = 0;
    coarsest_interval(dmax_f - dmin_f, width, &dmin, &range);
    sgm_pyramid* p = build_pyramid(width, height, num_layers, level_range, opt, dmin, range);
    if (p) {
        p->cam = *cam;   // level 0 uses the unscaled camera (CamParam_T::Scale(1))
        p->has_cam = true;
    }
    return p;
}

}  // extern "C"

namespace {

// Pipeline over the level-0 images in p->il[0] / p->ir[0], the result left
// in p->out[0]; or, with `io`, over the caller's device images io->img[0..1]
// with the result written to io->out_l (no staging copies).
int run_pyramid(sgm_pyramid* p, hipStream_t s, const IoBind* io = nullptr) {
    const int top = p->layers - 1;
    for (int i = 1; i <= top; ++i) {
        const uint8_t* srcl = (io && i == 1) ? io->img[0] : p->il[i - 1];
        const uint8_t* srcr = (io && i == 1) ? io->img[1] : p->ir[i - 1];
        PyrZoom z{srcl, srcr, p->il[i], p->ir[i], p->lw[i - 1], p->lh[i - 1]};
        PYR_TRY(pyr_launch_zoom(z, s));
    }
    StagePlan top_plan = p->top_plan;
    if (io && top == 0) top_plan.io = *io;
    PYR_TRY(launch_match(top_plan, false, s, nullptr));
    // the coarsest level of a multi-level pyramid is dilated
    // (FasterStereoCuda.cpp:192-197, cusgm_df.cu:300-318); its peaks are then
    // removed if asked, after the filters as in StereoCudaImpl::Match (:541-542)
    float* top_map = (io && top == 0) ? io->out_l : p->out[top];
    if (top > 0) {
        const int w = p->lw[top], h = p->lh[top];
        const int sm = std::min(w, h);
        PYR_TRY(pyr_launch_dilation(p->out[top], p->raw[top], w, h, sm < 360 ? 3 : (sm < 720 ? 5 : 7), s));
        top_map = p->raw[top];
    }
    if (p->opt.do_remove_peaks)
        PYR_TRY(launch_remove_peaks(top_map, p->peaks, p->lw[top], p->lh[top], peak_strips(),
                                    p->opt.peaks_ratio_threshold, s));
    for (int i = top - 1; i >= 0; --i) {
        PyrInit in{top > 0 && i == top - 1 ? p->raw[top] : p->out[i + 1], p->init[i], p->lw[i], p->lh[i], p->LR};
        PYR_TRY(pyr_launch_init_next(in, s));
        StagePlan lp = p->lvl_plan[i];
        if (io && i == 0) lp.io = *io;
        PYR_TRY(launch_census(lp, s));
        PyrLevel a{};
        a.census_l = p->cl[i]; a.census_r = p->cr[i]; a.init = p->init[i];
        a.cost = p->cost[i];
        for (int r = 0; r < 4; ++r) a.path[r] = p->path[i][r];
        a.disp = p->raw[i];
        a.w = p->lw[i]; a.h = p->lh[i]; a.level_range = p->LR;
        a.cs_type = p->opt.cs_type; a.p1 = p->opt.p1; a.p2 = p->opt.p2_init;
        a.uniqueness = (i == 0) ? p->opt.unique_threshold : 1.0f;
        PYR_TRY(pyr_launch_cost_init(a, s));
        PYR_TRY(pyr_launch_aggregate_init(a, s));
        PYR_TRY(pyr_launch_wta_init(a, s));
        PYR_TRY(launch_post(lp, s));
    }
    return SGM_OK;
}

}  // namespace

extern "C" {

int sgm_set_window_form(int32_t form) {
    if (form != SGM_WINDOW_AUTO && form != SGM_WINDOW_U32 && form != SGM_WINDOW_PAIRS &&
        form != SGM_WINDOW_QUADS)
        return set_global_error(SGM_E_INVALID_ARG,
                                "window form must be SGM_WINDOW_AUTO, _U32, _PAIRS or _QUADS");
    fsgm::set_window_form(form);
    return SGM_OK;
}

int32_t sgm_get_window_form(void) { return fsgm::window_form(); }

int sgm_pyramid_match_device(sgm_pyramid* p, const uint8_t* d_left, const uint8_t* d_right,
                             float* d_disp, void* stream) {
    if (!p || !d_left || !d_right || !d_disp)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: NULL argument");
    PYR_TRY(hipSetDevice(p->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the caller's buffers are bound in place of the level-0 image and map
    // buffers: nothing is staged
    IoBind io{};
    io.img[0] = d_left;
    io.img[1] = d_right;
    io.out_l = d_disp;
    io.out_r = nullptr;
    io.on = 1;
    return run_pyramid(p, s, &io);
}

int sgm_pyramid_match2_device(sgm_pyramid* p, const uint8_t* d_left, const uint8_t* d_right,
                              float* d_depth, void* stream) {
    if (!p || !d_left || !d_right || !d_depth)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: NULL argument");
    if (!p->has_cam)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: Match2 needs a pyramid from Init2");
    PYR_TRY(hipSetDevice(p->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t npix0 = (size_t)p->W * p->H;
    IoBind io{};
    io.img[0] = d_left;
    io.img[1] = d_right;
    io.out_l = p->out[0];
    io.out_r = nullptr;
    io.on = 1;
    if (int rc = run_pyramid(p, s, &io)) return rc;
    PYR_TRY(launch_depth(p->out[0], d_depth, npix0, p->cam.baseline * p->cam.focus,
                         p->cam.x0_right - p->cam.x0_left, s));
    return SGM_OK;
}

int sgm_pyramid_match2(sgm_pyramid* p, const uint8_t* left, const uint8_t* right, float* depth) {
    if (!p || !left || !right || !depth)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: NULL argument");
    if (!p->has_cam)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: Match2 needs a pyramid from Init2");
    PYR_TRY(hipSetDevice(p->device));
    const size_t npix0 = (size_t)p->W * p->H;
    PYR_TRY(hipMemcpyAsync(p->il[0], left, npix0, hipMemcpyHostToDevice, p->stream));
    PYR_TRY(hipMemcpyAsync(p->ir[0], right, npix0, hipMemcpyHostToDevice, p->stream));
    if (int rc = run_pyramid(p, p->stream)) return rc;
    // level 0's raw map is free once its post stage has read it
    PYR_TRY(launch_depth(p->out[0], p->raw[0], npix0, p->cam.baseline * p->cam.focus,
                         p->cam.x0_right - p->cam.x0_left, p->stream));
    PYR_TRY(hipMemcpyAsync(depth, p->raw[0], npix0 * 4, hipMemcpyDeviceToHost, p->stream));
    PYR_TRY(hipStreamSynchronize(p->stream));
    return SGM_OK;
}

int sgm_pyramid_match(sgm_pyramid* p, const uint8_t* left, const uint8_t* right, float* disp) {
    if (!p || !left || !right || !disp)
        return set_global_error(SGM_E_INVALID_ARG, "pyramid: NULL argument");
    PYR_TRY(hipSetDevice(p->device));
    const size_t npix0 = (size_t)p->W * p->H;
    PYR_TRY(hipMemcpyAsync(p->il[0], left, npix0, hipMemcpyHostToDevice, p->stream));
    PYR_TRY(hipMemcpyAsync(p->ir[0], right, npix0, hipMemcpyHostToDevice, p->stream));
    if (int rc = run_pyramid(p, p->stream)) return rc;
    PYR_TRY(hipMemcpyAsync(disp, p->out[0], npix0 * 4, hipMemcpyDeviceToHost, p->stream));
    PYR_TRY(hipStreamSynchronize(p->stream));
    return SGM_OK;
}

int32_t sgm_pyramid_coarsest_range(const sgm_pyramid* p, int32_t* min_disp) {
    if (!p) return 0;
    if (min_disp) *min_disp = p->top_min;
    return p->top_range;
}

size_t sgm_pyramid_device_bytes(const sgm_pyramid* p) { return p ? p->arena_size : 0; }

void sgm_pyramid_destroy(sgm_pyramid* p) {
    if (p) free_pyramid(p);
}

// ---- stage entry points of the finer levels (device pointers) ----
int sgm_stage_zoom(const uint8_t* d_src, int32_t w, int32_t h, uint8_t* d_dst, void* stream) {
    if (!d_src || !d_dst || w < 2 || h < 2) return set_global_error(SGM_E_INVALID_ARG, "zoom: bad args");
    PyrZoom z{d_src, d_src, d_dst, d_dst, w, h};
    hipStream_t s = static_cast<hipStream_t>(stream);
    PYR_TRY(pyr_launch_zoom(z, s));
    PYR_TRY(hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_stage_init_next(const float* d_coarse, int32_t w, int32_t h, int32_t level_range,
                        int16_t* d_next, void* stream) {
    if (!d_coarse || !d_next || w < 2 || h < 2 || (level_range != 16 && level_range != 32))
        return set_global_error(SGM_E_INVALID_ARG, "init_next: bad args");
    PyrInit a{d_coarse, d_next, w, h, level_range};
    hipStream_t s = static_cast<hipStream_t>(stream);
    PYR_TRY(pyr_launch_init_next(a, s));
    PYR_TRY(hipStreamSynchronize(s));
    return SGM_OK;
}

static int level_args(const uint64_t* cl, const uint64_t* cr, const int16_t* init, int32_t w,
                      int32_t h, int32_t lr, PyrLevel* a) {
    if (w <= 0 || h <= 0 || (lr != 16 && lr != 32) || !init)
        return set_global_error(SGM_E_INVALID_ARG, "window stage: bad args");
    *a = PyrLevel{};
    a->census_l = cl; a->census_r = cr; a->init = init;
    a->w = w; a->h = h; a->level_range = lr;
    return SGM_OK;
}

int sgm_stage_cost_init(const uint64_t* d_census_l, const uint64_t* d_census_r,
                        const int16_t* d_init, int32_t w, int32_t h, int32_t level_range,
                        int32_t cs_type, uint8_t* d_cost, void* stream) {
    PyrLevel a;
    if (int rc = level_args(d_census_l, d_census_r, d_init, w, h, level_range, &a)) return rc;
    if (!d_census_l || !d_census_r || !d_cost) return set_global_error(SGM_E_INVALID_ARG, "NULL");
    a.cost = d_cost; a.cs_type = cs_type ? 1 : 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PYR_TRY(pyr_launch_cost_init(a, s));
    PYR_TRY(hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_stage_aggregate_init(const uint8_t* d_cost, const int16_t* d_init, int32_t w, int32_t h,
                             int32_t level_range, int32_t p1, int32_t p2,
                             uint8_t* const d_paths[4], void* stream) {
    PyrLevel a;
    if (int rc = level_args(nullptr, nullptr, d_init, w, h, level_range, &a)) return rc;
    if (!d_cost || !d_paths || p1 < 0 || p2 < 0) return set_global_error(SGM_E_INVALID_ARG, "bad args");
    a.cost = const_cast<uint8_t*>(d_cost);
    for (int r = 0; r < 4; ++r) a.path[r] = d_paths[r];
    a.p1 = p1; a.p2 = p2;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PYR_TRY(pyr_launch_aggregate_init(a, s));
    PYR_TRY(hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_stage_wta_init(uint8_t* const d_paths[4], const int16_t* d_init, int32_t w, int32_t h,
                       int32_t level_range, float uniqueness, float* d_disp, void* stream) {
    PyrLevel a;
    if (int rc = level_args(nullptr, nullptr, d_init, w, h, level_range, &a)) return rc;
    if (!d_paths || !d_disp) return set_global_error(SGM_E_INVALID_ARG, "NULL");
    for (int r = 0; r < 4; ++r) a.path[r] = d_paths[r];
    a.disp = d_disp; a.uniqueness = uniqueness;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PYR_TRY(pyr_launch_wta_init(a, s));
    PYR_TRY(hipStreamSynchronize(s));
    return SGM_OK;
}

}  // extern "C"
