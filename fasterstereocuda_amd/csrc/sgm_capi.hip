// C ABI of the SGM stereo engine: handle lifecycle, device memory, plans,
// the StereoCuda-style Init / Match pipeline and the stage entry points.
//
// Pipeline per Match (ethan-li-coding/FasterStereoCuda,
// SgmStereoCuda/StereoCudaImpl.cpp:507-546): census of both views, cost of
// both views, 4-direction aggregation of both views, WTA of both views, LR
// check + median.  The reference runs the right view after the left one on
// a shared cost buffer; here both views are jobs of the same launches.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "sgm_capi.h"
#include "pyr_internal.h"
#include "sgm_internal.h"

using namespace fsgm;


namespace {

thread_local std::string g_last_error;

int fail(std::string* dst, int code, const std::string& msg) {
    if (dst) *dst = msg;
    g_last_error = msg;
    return code;
}

#define HIP_TRY(dst, expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail((dst), SGM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace

namespace fsgm {
int set_global_error(int code, const std::string& msg) { return fail(nullptr, code, msg); }

namespace {
std::atomic<int> g_cost_mode{0};
std::atomic<int> g_wta_form{0};
std::atomic<int> g_wide_min_groups{-1};
std::atomic<int> g_peak_strips{1};
}  // namespace
int cost_mode() { return g_cost_mode.load(); }
int peak_strips() { return g_peak_strips.load(); }

// Horizontal scanline groups a plan needs before SGM_COST_AUTO picks the
// on-the-fly kernels: two waves per SIMD of the current device.  With fewer,
// their long scanlines (1024 / D lines per wave) leave the GPU idle and the
// stored-cost kernels (4 lines per wave) finish first.
int wide_min_groups() {
    int v = g_wide_min_groups.load();
    if (v < 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            cus = 256;
        v = 8 * cus;
        g_wide_min_groups.store(v);
    }
    return v;
}

// Stored-cost plans: the four directions and then the WTA, or bottom->top
// fused with the WTA after the other three.  The fused bottom->top walk is
// latency-bound on a small plan (about 250 ns per step), so when the
// four-direction launch leaves the device at 2.5 waves per SIMD or fewer the
// walk is better run beside the horizontal ones and the WTA left to its own
// memory-bound launch -- unless the rows are much longer than the columns:
// then the fused walk is short and the extra launch costs more than it saves
// (measured on single pairs from 450x375 to 1920x1080).
bool separate_wta(const StagePlan& p) {
    const int f = g_wta_form.load();
    if (f == SGM_WTA_FUSED) return false;
    if (f == SGM_WTA_SEPARATE) return true;
    return p.squarish && 4 * p.n_agg <= 5 * wide_min_groups();
}

bool squarish_jobs(const std::vector<SgmJob>& jobs, int n) {
    for (int i = 0; i < n; ++i)
        if (2 * (int64_t)(jobs[i].W - jobs[i].roi_x) > 5 * (int64_t)jobs[i].H) return false;
    return n > 0;
}

bool choose_wide(bool eligible, int n_wh) {
    const int mode = cost_mode();
    if (!eligible || mode == SGM_COST_STORED) return false;
    return mode == SGM_COST_ON_THE_FLY || n_wh >= wide_min_groups();
}
}  // namespace fsgm

namespace {

inline size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

bool valid_range(int D) { return D == 64 || D == 128 || D == 256 || D == 512; }

// Validates the option fields the engine reads; returns SGM_OK or an error.
// allow_peaks: the plain handle removes peaks; the batch does not
int check_option(const sgm_option* o, std::string* err, bool allow_peaks) {
    if (!o) return fail(err, SGM_E_INVALID_ARG, "option pointer is NULL");
    if (o->cs_type != 0 && o->cs_type != 1)
        return fail(err, SGM_E_INVALID_ARG, "cs_type must be 0 (5x5) or 1 (9x7)");
    if (o->p1 < 0 || o->p2_init < 0 || (int64_t)o->p1 + o->p2_init > 65000)
        return fail(err, SGM_E_INVALID_ARG, "P1 and P2 must be >= 0 with P1 + P2 <= 65000");
    if (o->do_remove_peaks && !allow_peaks)
        return fail(err, SGM_E_UNSUPPORTED, "do_remove_peaks is not built for batches");
    if (o->do_remove_peaks && !(o->peaks_ratio_threshold >= 0.0f && o->peaks_ratio_threshold <= 1.0f))
        return fail(err, SGM_E_INVALID_ARG, "peaks_ratio_threshold must be in [0, 1]");
    if (o->post_filter_type != 0)
        return fail(err, SGM_E_UNSUPPORTED, "post filters (Gauss/bilateral) are not built");
    if (o->morphology_type != 0)
        return fail(err, SGM_E_UNSUPPORTED, "morphology filters are not built");
    // using_constant_p2 is not checked: the reference reads it only when it
    // aggregates with initial disparities (cusgm_ca.cu:4005-4008); the plain
    // aggregation ignores it (cusgm_ca.cu:4080-4135), and so does this one
    return SGM_OK;
}

}  // namespace

// Uploads jobs / posts and the per-stage item tables into one device block.
int fsgm::upload_plan(const std::vector<SgmJob>& jobs, const std::vector<SgmPost>& posts,
                StagePlan* plan, void** block, std::string* err) {
    const int nj = (int)jobs.size(), np = (int)posts.size();
    std::vector<int> oc(nj + 1, 0), ok(nj + 1, 0), oa(nj + 1, 0), ow(nj + 1, 0), op(np + 1, 0);
    std::vector<int> o3(nj + 1, 0), od(nj + 1, 0), w3(nj + 1, 0), wd(nj + 1, 0), wf(nj + 1, 0);
    int dmax = 64;
    bool wide = true;
    for (int i = 0; i < nj; ++i) {
        wide = wide && wide_ok(jobs[i]);
        const bool w = jobs[i].D <= 256;
        w3[i + 1] = w3[i] + (w ? wide_h_items(jobs[i]) : 0);
        wd[i + 1] = wd[i] + (w ? wide_v_items(jobs[i], kWideV) : 0);
        wf[i + 1] = wf[i] + (w ? wide_v_items(jobs[i], kWideF) : 0);
        oc[i + 1] = oc[i] + census_items(jobs[i]);
        ok[i + 1] = ok[i] + cost_items(jobs[i]);
        oa[i + 1] = oa[i] + agg_items(jobs[i]);
        ow[i + 1] = ow[i] + wta_items(jobs[i]);
        o3[i + 1] = o3[i] + agg3_items(jobs[i]);
        od[i + 1] = od[i] + dw_items(jobs[i]);
        dmax = std::max(dmax, jobs[i].D);
    }
    for (int i = 0; i < np; ++i) op[i + 1] = op[i] + post_items(posts[i]);
    const size_t s_jobs = align256(sizeof(SgmJob) * std::max(nj, 1));
    const size_t s_posts = align256(sizeof(SgmPost) * std::max(np, 1));
    const size_t s_oj = align256(sizeof(int) * (nj + 1));
    const size_t s_op = align256(sizeof(int) * (np + 1));
    const size_t total = s_jobs + s_posts + 9 * s_oj + s_op;
    std::vector<uint8_t> host(total, 0);
    size_t off = 0;
    auto put = [&](const void* src, size_t n, size_t slot) {
        std::memcpy(host.data() + off, src, n);
        const size_t at = off;
        off += slot;
        return at;
    };
    const size_t a_jobs = put(jobs.data(), sizeof(SgmJob) * nj, s_jobs);
    const size_t a_posts = put(posts.data(), sizeof(SgmPost) * np, s_posts);
    const size_t a_oc = put(oc.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_ok = put(ok.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_oa = put(oa.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_ow = put(ow.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_o3 = put(o3.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_od = put(od.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_w3 = put(w3.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_wd = put(wd.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_wf = put(wf.data(), sizeof(int) * (nj + 1), s_oj);
    const size_t a_op = put(op.data(), sizeof(int) * (np + 1), s_op);
    if (*block) {
        (void)hipFree(*block);
        *block = nullptr;
    }
    HIP_TRY(err, hipMalloc(block, total));
    HIP_TRY(err, hipMemcpy(*block, host.data(), total, hipMemcpyHostToDevice));
    uint8_t* b = static_cast<uint8_t*>(*block);
    plan->jobs = reinterpret_cast<SgmJob*>(b + a_jobs);
    plan->posts = reinterpret_cast<SgmPost*>(b + a_posts);
    plan->offs_census = reinterpret_cast<int*>(b + a_oc);
    plan->offs_cost = reinterpret_cast<int*>(b + a_ok);
    plan->offs_agg = reinterpret_cast<int*>(b + a_oa);
    plan->offs_wta = reinterpret_cast<int*>(b + a_ow);
    plan->offs_agg3 = reinterpret_cast<int*>(b + a_o3);
    plan->offs_dw = reinterpret_cast<int*>(b + a_od);
    plan->offs_wh = reinterpret_cast<int*>(b + a_w3);
    plan->offs_wv = reinterpret_cast<int*>(b + a_wd);
    plan->offs_wf = reinterpret_cast<int*>(b + a_wf);
    plan->offs_post = reinterpret_cast<int*>(b + a_op);
    plan->njobs = nj;
    plan->nposts = np;
    plan->n_census = oc[nj];
    plan->n_cost = ok[nj];
    plan->n_agg = oa[nj];
    plan->n_wta = ow[nj];
    plan->n_agg3 = o3[nj];
    plan->n_dw = od[nj];
    plan->n_wh = w3[nj];
    plan->n_wv = wd[nj];
    plan->n_wf = wf[nj];
    plan->wide = choose_wide(wide, w3[nj]);
    plan->squarish = squarish_jobs(jobs, nj);
    plan->n_post = op[np];
    plan->dmax = dmax;
    return SGM_OK;
}

// Runs cost .. WTA only for the first n jobs (the left views): their item
// offsets come first, so find_job over n jobs stays valid.
void fsgm::restrict_matching(StagePlan* P, const std::vector<SgmJob>& jobs, int n) {
    P->njobs_ctw = n;
    P->n_cost = P->n_agg = P->n_wta = P->n_agg3 = P->n_dw = P->n_wh = P->n_wv = P->n_wf = 0;
    bool wide = true;
    for (int i = 0; i < n; ++i) {
        wide = wide && wide_ok(jobs[i]);
        if (jobs[i].D <= 256) {
            P->n_wh += wide_h_items(jobs[i]);
            P->n_wv += wide_v_items(jobs[i], kWideV);
            P->n_wf += wide_v_items(jobs[i], kWideF);
        }
        P->n_cost += cost_items(jobs[i]);
        P->n_agg += agg_items(jobs[i]);
        P->n_wta += wta_items(jobs[i]);
        P->n_agg3 += agg3_items(jobs[i]);
        P->n_dw += dw_items(jobs[i]);
    }
    P->wide = choose_wide(wide, P->n_wh);
    P->squarish = squarish_jobs(jobs, n);
}

// census, cost, three directions, bottom->top + WTA (or the four directions,
// then the WTA: separate_wta), post, [right-map LR invalidation]; marks[k]
// (may be null) are recorded before stage k and after the last one.
hipError_t fsgm::launch_match(const StagePlan& plan, bool right_invalidation, hipStream_t s,
                              hipEvent_t* marks, const AggSchedule* sched) {
    StagePlan ctw = plan;
    if (plan.njobs_ctw > 0) ctw.njobs = plan.njobs_ctw;
    hipError_t e;
    auto mark = [&](int k) { return marks ? hipEventRecord(marks[k], s) : hipSuccess; };
    if ((e = mark(0)) != hipSuccess) return e;
    if ((e = launch_census(plan, s)) != hipSuccess) return e;
    if ((e = mark(1)) != hipSuccess) return e;
    // with the on-the-fly cost there is no cost volume to build
    if (!plan.wide && (e = launch_cost(ctw, s)) != hipSuccess) return e;
    if ((e = mark(2)) != hipSuccess) return e;
    const bool separate = !plan.wide && !sched && separate_wta(ctw);
    if (plan.wide)
        e = sched ? launch_wide3_sched(ctw, *sched, s) : launch_wide3(ctw, s);
    else if (separate)
        e = launch_aggregate(ctw, s);
    else
        e = sched ? launch_aggregate_sched(ctw, *sched, s) : launch_aggregate3(ctw, s);
    if (e != hipSuccess) return e;
    if ((e = mark(3)) != hipSuccess) return e;
    if (plan.wide)
        e = sched ? launch_wide_wta_sched(ctw, *sched, s) : launch_wide_wta(ctw, s);
    else if (separate)
        e = launch_wta(ctw, s);
    else
        e = sched ? launch_aggregate_wta_sched(ctw, *sched, s) : launch_aggregate_wta(ctw, s);
    if (e != hipSuccess) return e;
    if ((e = mark(4)) != hipSuccess) return e;
    if ((e = launch_post(plan, s)) != hipSuccess) return e;
    if (right_invalidation && (e = launch_right_invalidate(plan, s)) != hipSuccess) return e;
    return mark(5);
}

// Per-view device buffers of one pair.
struct ViewBuffers {
    uint8_t* img = nullptr;
    uint64_t* census = nullptr;
    uint8_t* cost = nullptr;
    uint8_t* path[4] = {nullptr, nullptr, nullptr, nullptr};
    float* raw = nullptr;
    float* out = nullptr;
};

struct sgm_handle {
    bool inited = false;
    int device = -1;
    int32_t W = 0, H = 0, min_d = 0, D = 0;
    // D = LEVEL_RANGE (32): the handle matches around an initial disparity
    // map only (window cost, window aggregation, window WTA), as the
    // reference's finer pyramid levels do (FasterStereoCuda.cpp:282-289, 455-469)
    bool window = false;
    int16_t* init = nullptr;      // W x H initial disparities (window handles)
    sgm_option opt{};
    bool print_log = false;
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    void* plan_block = nullptr;
    ViewBuffers view[2];
    StagePlan plan;
    bool has_cam = false;         // set by sgm_init2, used by sgm_match2
    sgm_camera cam{};
    int* peaks = nullptr;         // 2 W H ints when the option removes peaks
    std::string err;
};

namespace {

size_t arena_bytes(int W, int H, int D) {
    const size_t npix = (size_t)W * H, ncell = npix * D;
    return 2 * (align256(npix) + align256(npix * 8) + 5 * align256(ncell) + 2 * align256(npix * 4)) +
           align256(npix * 2);
}

void carve(sgm_handle* h) {
    const size_t npix = (size_t)h->W * h->H, ncell = npix * h->D;
    uint8_t* p = static_cast<uint8_t*>(h->arena);
    auto take = [&](size_t n) { uint8_t* r = p; p += align256(n); return r; };
    for (int v = 0; v < 2; ++v) {
        ViewBuffers& b = h->view[v];
        b.img = take(npix);
        b.census = reinterpret_cast<uint64_t*>(take(npix * 8));
        b.cost = take(ncell);
        for (int r = 0; r < 4; ++r) b.path[r] = take(ncell);
        b.raw = reinterpret_cast<float*>(take(npix * 4));
        b.out = reinterpret_cast<float*>(take(npix * 4));
    }
    h->init = reinterpret_cast<int16_t*>(take(npix * 2));
}

SgmJob make_job(const sgm_handle* h, int v) {
    const ViewBuffers& b = h->view[v];
    SgmJob j{};
    j.img = b.img;
    j.census = b.census;
    j.census_other = h->view[1 - v].census;
    j.cost = b.cost;
    for (int r = 0; r < 4; ++r) j.path[r] = b.path[r];
    j.disp = b.raw;
    j.W = h->W; j.H = h->H; j.D = h->D; j.min_d = h->min_d;
    j.roi_x = std::max(0, h->min_d);
    j.is_right = v;
    j.cs_type = h->opt.cs_type;
    j.p1 = h->opt.p1; j.p2 = h->opt.p2_init;
    j.uniqueness = h->opt.unique_threshold;
    return j;
}

int build_handle_plan(sgm_handle* h) {
    std::vector<SgmJob> jobs;
    jobs.push_back(make_job(h, 0));
    jobs.push_back(make_job(h, 1));  // the right census is always needed
    SgmPost p{};
    p.raw_l = h->view[0].raw;
    p.raw_r = h->view[1].raw;
    p.out_l = h->view[0].out;
    p.out_r = h->view[1].out;
    p.W = h->W; p.H = h->H;
    p.roi_x = std::max(0, h->min_d);
    p.do_lr = h->opt.do_lr_check ? 1 : 0;
    p.do_median = h->opt.do_median_filter ? 1 : 0;
    p.lr_thr = h->opt.lr_threshold;
    if (h->window) {
        // census of both views, then the median of the window WTA's map; the
        // window pass covers every column and has no LR check
        for (SgmJob& j : jobs) { j.D = 64; j.min_d = 0; j.roi_x = 0; }
        p.roi_x = 0;
        p.do_lr = 0;
    }
    std::vector<SgmPost> posts{p};
    if (int rc = upload_plan(jobs, posts, &h->plan, &h->plan_block, &h->err)) return rc;
    if (!h->opt.do_lr_check && !h->window) restrict_matching(&h->plan, jobs, 1);
    return SGM_OK;
}

// The window pass of a D = LEVEL_RANGE handle (StereoCudaImpl::Match with
// init_disp_left, StereoCudaImpl.cpp:363-446; cusgm_cc.cu:173-237,
// cusgm_ca.cu:1181-1339, 3425-3512): census of both views, window cost,
// window aggregation, window WTA, median; marks as in launch_match.
hipError_t launch_window(const sgm_handle* h, const StagePlan& plan, hipStream_t s,
                         hipEvent_t* marks) {
    auto mark = [&](int k) { return marks ? hipEventRecord(marks[k], s) : hipSuccess; };
    PyrLevel a{};
    a.census_l = h->view[0].census; a.census_r = h->view[1].census; a.init = h->init;
    a.cost = h->view[0].cost;
    for (int r = 0; r < 4; ++r) a.path[r] = h->view[0].path[r];
    a.disp = h->view[0].raw;
    a.w = h->W; a.h = h->H; a.level_range = h->D;
    a.cs_type = h->opt.cs_type; a.p1 = h->opt.p1; a.p2 = h->opt.p2_init;
    a.uniqueness = h->opt.unique_threshold;
    hipError_t e;
    if ((e = mark(0)) != hipSuccess || (e = launch_census(plan, s)) != hipSuccess) return e;
    if ((e = mark(1)) != hipSuccess || (e = pyr_launch_cost_init(a, s)) != hipSuccess) return e;
    if ((e = mark(2)) != hipSuccess || (e = pyr_launch_aggregate_init(a, s)) != hipSuccess) return e;
    if ((e = mark(3)) != hipSuccess || (e = pyr_launch_wta_init(a, s)) != hipSuccess) return e;
    if ((e = mark(4)) != hipSuccess || (e = launch_post(plan, s)) != hipSuccess) return e;
    return mark(5);
}

// Runs the pipeline; `io` (may be null) binds caller buffers in place of the
// handle's images and outputs.  With the LR check off the right view is still
// censused (the left cost needs it) but not aggregated.
int run_pipeline(sgm_handle* h, bool want_right, hipStream_t s, const IoBind* io = nullptr) {
    const bool lr = h->opt.do_lr_check != 0;
    StagePlan plan = h->plan;
    if (io) plan.io = *io;
    float* const final_left = io ? io->out_l : h->view[0].out;
    // peak removal on the final left map, as StereoCudaImpl::Match does after
    // the filters (StereoCudaImpl.cpp:541-542)
    auto peaks = [&]() {
        return h->opt.do_remove_peaks
                   ? launch_remove_peaks(final_left, h->peaks, h->W, h->H, peak_strips(),
                                         h->opt.peaks_ratio_threshold, s)
                   : hipSuccess;
    };
    auto run = [&](hipEvent_t* marks) {
        return h->window ? launch_window(h, plan, s, marks)
                         : launch_match(plan, want_right && lr, s, marks);
    };
    if (!h->print_log) {
        HIP_TRY(&h->err, run(nullptr));
        HIP_TRY(&h->err, peaks());
        return SGM_OK;
    }
    hipEvent_t ev[6];
    for (auto& e : ev) HIP_TRY(&h->err, hipEventCreate(&e));
    HIP_TRY(&h->err, run(ev));
    HIP_TRY(&h->err, peaks());
    HIP_TRY(&h->err, hipEventSynchronize(ev[5]));
    static const char* names[5] = {"Census:", "Cost:", "Aggregate (3 paths):",
                                   "Aggregate + WTA:", "LRCheck/Filter:"};
    static const char* wnames[5] = {"Census:", "Cost (window):", "Aggregate (window):",
                                    "WTA (window):", "Filter:"};
    for (int k = 0; k < 5; ++k) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
        std::printf("** %-24s %.3f ms\n", h->window ? wnames[k] : names[k], ms);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return SGM_OK;
}

int check_ready(sgm_handle* h) {
    if (!h) return fail(nullptr, SGM_E_INVALID_ARG, "handle is NULL");
    if (!h->inited) return fail(&h->err, SGM_E_NOT_INIT, "handle is not initialised");
    HIP_TRY(&h->err, hipSetDevice(h->device));
    return SGM_OK;
}

// Which pass a match with (init != NULL) or without an initial disparity map
// runs on this handle, or the error that refuses it.
int check_match_kind(sgm_handle* h, const void* init, const sgm_roi* rl, const sgm_roi* rr) {
    if ((rl || rr) && init)
        return fail(&h->err, SGM_E_UNSUPPORTED,
                    "regions of interest with an initial disparity map are not built");
    if ((rl && (rl->w < 0 || rl->h < 0)) || (rr && (rr->w < 0 || rr->h < 0)))
        return fail(&h->err, SGM_E_INVALID_ARG, "a region of interest has a negative size");
    if (!init) {
        if (h->window)
            return fail(&h->err, SGM_E_UNSUPPORTED,
                        "a disp_range 32 handle matches with an initial disparity map only");
        return SGM_OK;
    }
    if (!h->window)
        return fail(&h->err, SGM_E_UNSUPPORTED,
                    "initial disparities need disp_range 32 (LEVEL_RANGE); the reference's "
                    "D = 64 kernels with initial disparities are not built");
    if (h->opt.do_lr_check)
        return fail(&h->err, SGM_E_UNSUPPORTED,
                    "the LR check with initial disparities is not built");
    if (!h->opt.using_constant_p2)
        return fail(&h->err, SGM_E_UNSUPPORTED,
                    "with initial disparities the reference aggregates only with constant P2");
    if (h->min_d > 0)
        return fail(&h->err, SGM_E_UNSUPPORTED,
                    "initial disparities with min_disparity > 0 (a column offset) are not built");
    return SGM_OK;
}

// the plain pass with regions of interest (defined with the stage helpers)
int run_roi(sgm_handle* h, const sgm_roi* rl, const sgm_roi* rr, bool want_right, hipStream_t s);

}  // namespace

extern "C" {

sgm_handle* sgm_create(void) { return new (std::nothrow) sgm_handle(); }

void sgm_destroy(sgm_handle* h) {
    if (!h) return;
    sgm_release(h);
    delete h;
}

int sgm_init(sgm_handle* h, int32_t width, int32_t height, int32_t min_disparity,
             int32_t disp_range, const sgm_option* opt, int32_t print_log) {
    if (!h) return fail(nullptr, SGM_E_INVALID_ARG, "handle is NULL");
    // validate everything before touching the device
    if (width <= 0 || height <= 0)
        return fail(&h->err, SGM_E_INVALID_ARG, "image size must be positive");
    if (!valid_range(disp_range) && disp_range != SGM_LEVEL_RANGE)
        return fail(&h->err, SGM_E_INVALID_ARG,
                    "disp_range must be 64, 128, 256 or 512 (or 32 = LEVEL_RANGE for matching "
                    "around initial disparities)");
    if (std::max(0, min_disparity) >= width)
        return fail(&h->err, SGM_E_INVALID_ARG, "min_disparity leaves no column to match");
    if (int rc = check_option(opt, &h->err, true)) return rc;
    if (opt->do_remove_peaks && (int64_t)width * height > INT32_MAX)
        return fail(&h->err, SGM_E_INVALID_ARG, "peak removal needs width * height < 2^31");
    sgm_release(h);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(&h->err, SGM_E_NO_DEVICE, "no HIP device");
    HIP_TRY(&h->err, hipGetDevice(&h->device));
    h->W = width; h->H = height; h->min_d = min_disparity; h->D = disp_range;
    h->window = disp_range == SGM_LEVEL_RANGE;
    h->opt = *opt;
    h->print_log = print_log != 0;
    HIP_TRY(&h->err, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIP_TRY(&h->err, hipMalloc(&h->arena, arena_bytes(width, height, disp_range)));
    if (opt->do_remove_peaks)
        HIP_TRY(&h->err, hipMalloc(&h->peaks, 2 * (size_t)width * height * sizeof(int)));
    carve(h);
    if (int rc = build_handle_plan(h)) return rc;
    h->inited = true;
    if (h->print_log)
        std::printf("Init Completed., Disp: %d~%d\n", min_disparity, min_disparity + disp_range);
    return SGM_OK;
}

int sgm_match_device(sgm_handle* h, const uint8_t* d_left, const uint8_t* d_right,
                     float* d_disp_left, float* d_disp_right, void* stream) {
    if (int rc = check_ready(h)) return rc;
    if (!d_left || !d_right || !d_disp_left)
        return fail(&h->err, SGM_E_INVALID_ARG, "image or output pointer is NULL");
    if (int rc = check_match_kind(h, nullptr, nullptr, nullptr)) return rc;
    // the caller's buffers are read and written in place: no staging copies
    IoBind io{};
    io.img[0] = d_left;
    io.img[1] = d_right;
    io.out_l = d_disp_left;
    io.out_r = d_disp_right;
    io.on = 1;
    return run_pipeline(h, d_disp_right != nullptr, static_cast<hipStream_t>(stream), &io);
}

int sgm_match(sgm_handle* h, const uint8_t* img_left, const uint8_t* img_right,
              float* disp_left, float* disp_right, const int16_t* init_disp_left,
              const sgm_roi* roi_left, const sgm_roi* roi_right) {
    if (int rc = check_ready(h)) return rc;
    if (!img_left || !img_right || !disp_left)
        return fail(&h->err, SGM_E_INVALID_ARG, "image or output pointer is NULL");
    if (int rc = check_match_kind(h, init_disp_left, roi_left, roi_right)) return rc;
    const size_t npix = (size_t)h->W * h->H;
    hipStream_t s = h->stream;
    HIP_TRY(&h->err, hipMemcpyAsync(h->view[0].img, img_left, npix, hipMemcpyHostToDevice, s));
    HIP_TRY(&h->err, hipMemcpyAsync(h->view[1].img, img_right, npix, hipMemcpyHostToDevice, s));
    if (init_disp_left)
        HIP_TRY(&h->err, hipMemcpyAsync(h->init, init_disp_left, npix * 2, hipMemcpyHostToDevice, s));
    if (roi_left || roi_right) {
        if (int rc = run_roi(h, roi_left, roi_right, disp_right != nullptr, s)) return rc;
    } else if (int rc = run_pipeline(h, disp_right != nullptr, s)) {
        return rc;
    }
    HIP_TRY(&h->err, hipMemcpyAsync(disp_left, h->view[0].out, npix * 4, hipMemcpyDeviceToHost, s));
    if (disp_right)
        HIP_TRY(&h->err, hipMemcpyAsync(disp_right, h->view[1].out, npix * 4,
                                        hipMemcpyDeviceToHost, s));
    HIP_TRY(&h->err, hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_depth_device(const float* d_disp, int32_t width, int32_t height, const sgm_camera* cam,
                     float* d_depth, void* stream) {
    if (!d_disp || !d_depth || !cam || width <= 0 || height <= 0)
        return fail(nullptr, SGM_E_INVALID_ARG, "depth: NULL pointer or bad size");
    HIP_TRY(nullptr, launch_depth(d_disp, d_depth, (size_t)width * height,
                                  cam->baseline * cam->focus, cam->x0_right - cam->x0_left,
                                  static_cast<hipStream_t>(stream)));
    return SGM_OK;
}

int sgm_init2(sgm_handle* h, int32_t width, int32_t height, int32_t min_disparity,
              int32_t disp_range, const sgm_option* opt, const sgm_camera* cam, int32_t print_log) {
    if (!h) return fail(nullptr, SGM_E_INVALID_ARG, "handle is NULL");
    if (!cam) return fail(&h->err, SGM_E_INVALID_ARG, "camera is NULL");
    if (int rc = sgm_init(h, width, height, min_disparity, disp_range, opt, print_log)) return rc;
    h->cam = *cam;
    h->has_cam = true;
    return SGM_OK;
}

int sgm_match2(sgm_handle* h, const uint8_t* img_left, const uint8_t* img_right, float* depth_left,
               const int16_t* init_disp_left, const sgm_roi* roi_left, const sgm_roi* roi_right) {
    if (int rc = check_ready(h)) return rc;
    if (!h->has_cam) return fail(&h->err, SGM_E_INVALID_ARG, "Match2 needs a handle set up by Init2");
    if (!img_left || !img_right || !depth_left)
        return fail(&h->err, SGM_E_INVALID_ARG, "image or output pointer is NULL");
    if (int rc = check_match_kind(h, init_disp_left, roi_left, roi_right)) return rc;
    const size_t npix = (size_t)h->W * h->H;
    hipStream_t s = h->stream;
    HIP_TRY(&h->err, hipMemcpyAsync(h->view[0].img, img_left, npix, hipMemcpyHostToDevice, s));
    HIP_TRY(&h->err, hipMemcpyAsync(h->view[1].img, img_right, npix, hipMemcpyHostToDevice, s));
    if (init_disp_left)
        HIP_TRY(&h->err, hipMemcpyAsync(h->init, init_disp_left, npix * 2, hipMemcpyHostToDevice, s));
    if (roi_left || roi_right) {
        if (int rc = run_roi(h, roi_left, roi_right, false, s)) return rc;
    } else if (int rc = run_pipeline(h, false, s)) {
        return rc;
    }
    // the raw left map is free once the post stage has read it
    float* depth = h->view[0].raw;
    HIP_TRY(&h->err, launch_depth(h->view[0].out, depth, npix, h->cam.baseline * h->cam.focus,
                                  h->cam.x0_right - h->cam.x0_left, s));
    HIP_TRY(&h->err, hipMemcpyAsync(depth_left, depth, npix * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(&h->err, hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_set_min_disparity(sgm_handle* h, int32_t min_disparity) {
    if (int rc = check_ready(h)) return rc;
    if (std::max(0, min_disparity) >= h->W)
        return fail(&h->err, SGM_E_INVALID_ARG, "min_disparity leaves no column to match");
    h->min_d = min_disparity;
    return build_handle_plan(h);
}

int32_t sgm_get_disparity_range(const sgm_handle* h) { return h ? h->D : 0; }

int sgm_release(sgm_handle* h) {
    if (!h) return fail(nullptr, SGM_E_INVALID_ARG, "handle is NULL");
    if (h->device >= 0) (void)hipSetDevice(h->device);
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
        h->stream = nullptr;
    }
    if (h->arena) { (void)hipFree(h->arena); h->arena = nullptr; }
    if (h->plan_block) { (void)hipFree(h->plan_block); h->plan_block = nullptr; }
    if (h->peaks) { (void)hipFree(h->peaks); h->peaks = nullptr; }
    for (auto& v : h->view) v = ViewBuffers{};
    h->plan = StagePlan{};
    h->inited = false;
    h->has_cam = false;
    h->window = false;
    h->init = nullptr;
    return SGM_OK;
}

const char* sgm_last_error(const sgm_handle* h) {
    if (h) return h->err.c_str();
    return g_last_error.c_str();
}

int sgm_host_alloc(void** ptr, size_t size) {
    if (!ptr) return fail(nullptr, SGM_E_INVALID_ARG, "ptr is NULL");
    HIP_TRY(nullptr, hipHostMalloc(ptr, size, hipHostMallocDefault));
    return SGM_OK;
}

int sgm_host_free(void* ptr) {
    if (!ptr) return SGM_OK;
    HIP_TRY(nullptr, hipHostFree(ptr));
    return SGM_OK;
}

// StereoCudaImpl::GetRoiFromDispMap (StereoCudaImpl.cpp:585-628): first and
// last valid row by a raster scan, first and last valid column by a column
// scan.
// A map with no valid pixel leaves x and w as they were, as the reference's
// scan does; an empty size leaves the whole ROI untouched.
void sgm_roi_from_disparity(const float* disp, int32_t width, int32_t height, sgm_roi* roi) {
    if (!disp || !roi || width <= 0 || height <= 0) return;
    const int64_t size = (int64_t)width * height;
    int64_t n = 0;
    for (n = 0; n < size; ++n)
        if (disp[n] != SGM_INVALID_FLOAT) break;
    roi->y = (int32_t)(n / width);
    const int64_t size0 = (int64_t)roi->y * width;
    for (n = size - 1; n > size0; --n)
        if (disp[n] != SGM_INVALID_FLOAT) break;
    roi->h = (int32_t)(n / width) - roi->y + 1;
    bool done = false;
    for (int32_t j = 0; j < width && !done; ++j)
        for (int32_t i = 0; i < height; ++i)
            if (disp[(int64_t)i * width + j] != SGM_INVALID_FLOAT) { roi->x = j; done = true; break; }
    done = false;
    for (int32_t j = width - 1; j > roi->x && !done; --j)
        for (int32_t i = 0; i < height; ++i)
            if (disp[(int64_t)i * width + j] != SGM_INVALID_FLOAT) {
                roi->w = j - roi->x + 1; done = true; break;
            }
}

int sgm_set_cost_mode(int32_t mode) {
    if (mode != SGM_COST_AUTO && mode != SGM_COST_STORED && mode != SGM_COST_ON_THE_FLY)
        return fail(nullptr, SGM_E_INVALID_ARG,
                    "cost mode must be SGM_COST_AUTO, SGM_COST_STORED or SGM_COST_ON_THE_FLY");
    fsgm::g_cost_mode.store(mode);
    return SGM_OK;
}

int32_t sgm_get_cost_mode(void) { return fsgm::cost_mode(); }

int sgm_set_wta_form(int32_t form) {
    if (form != SGM_WTA_AUTO && form != SGM_WTA_FUSED && form != SGM_WTA_SEPARATE)
        return fail(nullptr, SGM_E_INVALID_ARG,
                    "WTA form must be SGM_WTA_AUTO, SGM_WTA_FUSED or SGM_WTA_SEPARATE");
    fsgm::g_wta_form.store(form);
    return SGM_OK;
}

int32_t sgm_get_wta_form(void) { return fsgm::g_wta_form.load(); }

int sgm_set_peak_strips(int32_t strips) {
    if (strips < 1 || strips > 4096)
        return fail(nullptr, SGM_E_INVALID_ARG, "peak strips must be in 1..4096");
    fsgm::g_peak_strips.store(strips);
    return SGM_OK;
}

int32_t sgm_get_peak_strips(void) { return fsgm::peak_strips(); }

float sgm_invalid_float(void) { return SGM_INVALID_FLOAT; }
int16_t sgm_invalid_short(void) { return SGM_INVALID_SHORT; }
int16_t sgm_level_range(void) { return SGM_LEVEL_RANGE; }

}  // extern "C"

// --------------------------------------------------------------- batches
struct sgm_batch {
    int device = -1;
    sgm_option opt{};
    void* arena = nullptr;
    size_t arena_size = 0;
    void* plan_block = nullptr;
    StagePlan plan;
    bool want_right = false;
    bool timing = false;
    std::vector<hipEvent_t> pool;     // free events
    std::vector<hipEvent_t> marks;    // 6 per recorded run
    std::string err;
    // scanline schedule
    std::vector<SgmJob> jobs;         // host copy of the matched jobs
    std::string signature;
    int policy = SGM_SCHED_GRID;
    int waves = 0;                    // resident workgroups of the persistent launches
    int per_cu = 0, n_cu = 0;
    void* sched_block = nullptr;      // items (agg3 then dw) + 2 counters
    AggSchedule sched;
    // graph mode
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

hipError_t take_event(sgm_batch* b, hipEvent_t* e) {
    if (!b->pool.empty()) {
        *e = b->pool.back();
        b->pool.pop_back();
        return hipSuccess;
    }
    return hipEventCreate(e);
}

void free_batch(sgm_batch* b) {
    if (b->device >= 0) (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->sched_block) (void)hipFree(b->sched_block);
    for (hipEvent_t e : b->marks) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->pool) (void)hipEventDestroy(e);
    if (b->arena) (void)hipFree(b->arena);
    if (b->plan_block) (void)hipFree(b->plan_block);
    delete b;
}


// Schedules remembered per batch signature (process-wide).
std::mutex g_sched_mu;
std::map<std::string, std::pair<int, int>> g_sched_cache;

std::string batch_signature(int device, const sgm_option& o, const sgm_pair_desc* pairs, int n,
                            bool want_right) {
    std::string s = "dev" + std::to_string(device) + "|cs" + std::to_string(o.cs_type) + "|p" +
                    std::to_string(o.p1) + "," + std::to_string(o.p2_init) + "|lr" +
                    std::to_string(o.do_lr_check) + "|r" + std::to_string(want_right ? 1 : 0);
    for (int i = 0; i < n; ++i)
        s += "|" + std::to_string(pairs[i].width) + "x" + std::to_string(pairs[i].height) + "d" +
             std::to_string(pairs[i].min_disparity) + "+" + std::to_string(pairs[i].disp_range);
    return s;
}

// Packed aggregation items (job << 16 | direction << 14 | group) of the
// matched jobs, in the order `policy` asks for.  Directions 0..3 are
// top->bottom, left->right, bottom->top, right->left.
int build_items(const std::vector<SgmJob>& jobs, bool wide, int policy, std::vector<uint32_t>* agg3,
                std::vector<uint32_t>* dw) {
    struct Item { int64_t work; uint32_t code; };
    std::vector<Item> a, v, b;   // horizontal, top->bottom, bottom->top
    for (size_t j = 0; j < jobs.size(); ++j) {
        const SgmJob& J = jobs[j];
        // scanlines per group of each kernel
        const int lh = wide ? wide_lines(kWideH, J.D) : 4, lv = wide ? wide_lines(kWideV, J.D) : 4;
        const int lf = wide ? wide_lines(kWideF, J.D) : 4;
        const int wv = J.W - J.roi_x;
        auto code = [&](int dir, int g) { return (uint32_t)j << 16 | (uint32_t)dir << 14 | (uint32_t)g; };
        for (int dir : {1, 3})   // left->right, right->left
            for (int g = 0; g < (J.H + lh - 1) / lh; ++g)
                a.push_back({(int64_t)std::min(lh, J.H - lh * g) * wv * J.D, code(dir, g)});
        for (int g = 0; g < (wv + lv - 1) / lv; ++g)
            v.push_back({(int64_t)std::min(lv, wv - lv * g) * J.H * J.D, code(0, g)});
        for (int g = 0; g < (wv + lf - 1) / lf; ++g)
            b.push_back({(int64_t)std::min(lf, wv - lf * g) * J.H * J.D, code(2, g)});
    }
    if (!wide) {   // one kernel takes every direction: a single list
        a.insert(a.end(), v.begin(), v.end());
        v.clear();
    }
    if (policy == SGM_SCHED_LONGEST_FIRST) {
        auto by_work = [](const Item& x, const Item& y) { return x.work > y.work; };
        std::stable_sort(a.begin(), a.end(), by_work);
        std::stable_sort(v.begin(), v.end(), by_work);
        std::stable_sort(b.begin(), b.end(), by_work);
    }
    agg3->clear();
    dw->clear();
    for (const Item& x : a) agg3->push_back(x.code);
    for (const Item& x : v) agg3->push_back(x.code);
    for (const Item& x : b) dw->push_back(x.code);
    return (int)a.size();
}

void drop_graph(sgm_batch* b) {
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
}

int apply_schedule(sgm_batch* b, int policy, int waves) {
    if (policy < SGM_SCHED_GRID || policy > SGM_SCHED_LONGEST_FIRST || waves < 0)
        return fail(&b->err, SGM_E_INVALID_ARG, "unknown schedule policy or negative wave count");
    drop_graph(b);
    if (policy == SGM_SCHED_GRID) {
        b->policy = policy;
        b->waves = 0;
        return SGM_OK;
    }
    if (b->jobs.size() > 65535)
        return fail(&b->err, SGM_E_INVALID_ARG, "persistent schedules take at most 65535 views");
    for (const SgmJob& J : b->jobs)
        if (J.H > 65532 || J.W > 65532)
            return fail(&b->err, SGM_E_INVALID_ARG, "persistent schedules take sides up to 65532");
    std::vector<uint32_t> a3, dw;
    const int n3h = build_items(b->jobs, b->plan.wide, policy, &a3, &dw);
    const size_t bytes = (a3.size() + dw.size() + 3) * sizeof(uint32_t);
    HIP_TRY(&b->err, hipSetDevice(b->device));
    if (!b->sched_block) HIP_TRY(&b->err, hipMalloc(&b->sched_block, bytes));
    std::vector<uint32_t> host(a3);
    host.insert(host.end(), dw.begin(), dw.end());
    host.push_back(0u);
    host.push_back(0u);
    host.push_back(0u);
    // earlier runs may still read the old list
    HIP_TRY(&b->err, hipDeviceSynchronize());
    HIP_TRY(&b->err, hipMemcpy(b->sched_block, host.data(), bytes, hipMemcpyHostToDevice));
    uint32_t* base = static_cast<uint32_t*>(b->sched_block);
    b->sched.items3 = base;
    b->sched.n3 = (int)a3.size();
    b->sched.n3h = n3h;
    b->sched.items_dw = base + a3.size();
    b->sched.ndw = (int)dw.size();
    b->sched.counters = base + a3.size() + dw.size();
    b->policy = policy;
    b->waves = waves > 0 ? waves : b->per_cu * b->n_cu;
    b->sched.grid = b->waves;
    return SGM_OK;
}

hipError_t enqueue(sgm_batch* b, hipStream_t s, hipEvent_t* marks) {
    return launch_match(b->plan, b->want_right && b->opt.do_lr_check, s, marks,
                        b->policy == SGM_SCHED_GRID ? nullptr : &b->sched);
}

int run_graph(sgm_batch* b, hipStream_t s) {
    if (!b->exec) {
        if (!b->cap_stream) HIP_TRY(&b->err, hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        HIP_TRY(&b->err, hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        const hipError_t le = enqueue(b, b->cap_stream, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(b->cap_stream, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return fail(&b->err, SGM_E_HIP, std::string("graph capture failed: ") +
                                                 hipGetErrorString(le != hipSuccess ? le : ce));
        }
        const hipError_t ie = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ie != hipSuccess) {
            b->exec = nullptr;
            return fail(&b->err, SGM_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
        }
    }
    HIP_TRY(&b->err, hipGraphLaunch(b->exec, s));
    return SGM_OK;
}
}  // namespace

extern "C" {

sgm_batch* sgm_batch_create(const sgm_pair_desc* pairs, int32_t n, const sgm_option* opt) {
    std::string err;
    if (!pairs || n <= 0) { fail(nullptr, SGM_E_INVALID_ARG, "empty batch"); return nullptr; }
    if (check_option(opt, &err, false)) return nullptr;
    for (int i = 0; i < n; ++i) {
        const sgm_pair_desc& p = pairs[i];
        if (p.width <= 0 || p.height <= 0 || !valid_range(p.disp_range) ||
            std::max(0, p.min_disparity) >= p.width) {
            fail(nullptr, SGM_E_INVALID_ARG, "pair " + std::to_string(i) + ": bad size or range");
            return nullptr;
        }
        if (!p.d_left || !p.d_right || !p.d_disp_left) {
            fail(nullptr, SGM_E_INVALID_ARG, "pair " + std::to_string(i) + ": NULL pointer");
            return nullptr;
        }
    }
    sgm_batch* b = new (std::nothrow) sgm_batch();
    if (!b) { fail(nullptr, SGM_E_INVALID_ARG, "out of host memory"); return nullptr; }
    b->opt = *opt;
    if (hipGetDevice(&b->device) != hipSuccess) {
        fail(nullptr, SGM_E_NO_DEVICE, "no HIP device");
        delete b;
        return nullptr;
    }
    const bool lr = opt->do_lr_check != 0;
    // arena: per pair two census maps; cost + 4 paths + raw map per matched view
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        const size_t npix = (size_t)pairs[i].width * pairs[i].height;
        const size_t ncell = npix * pairs[i].disp_range;
        total += 2 * align256(npix * 8) + (lr ? 2 : 1) * (5 * align256(ncell) + align256(npix * 4));
    }
    if (hipMalloc(&b->arena, total) != hipSuccess) {
        fail(nullptr, SGM_E_HIP, "hipMalloc of " + std::to_string(total) + " bytes failed");
        free_batch(b);
        return nullptr;
    }
    b->arena_size = total;
    uint8_t* cur = static_cast<uint8_t*>(b->arena);
    auto take = [&](size_t bytes) { uint8_t* r = cur; cur += align256(bytes); return r; };
    std::vector<SgmJob> left_jobs, right_jobs;
    std::vector<SgmPost> posts;
    for (int i = 0; i < n; ++i) {
        const sgm_pair_desc& p = pairs[i];
        const size_t npix = (size_t)p.width * p.height, ncell = npix * p.disp_range;
        uint64_t* cs[2] = {reinterpret_cast<uint64_t*>(take(npix * 8)),
                           reinterpret_cast<uint64_t*>(take(npix * 8))};
        SgmPost post{};
        for (int v = 0; v < (lr ? 2 : 1); ++v) {
            SgmJob j{};
            j.img = v ? p.d_right : p.d_left;
            j.census = cs[v];
            j.census_other = cs[1 - v];
            j.cost = take(ncell);
            for (int r = 0; r < 4; ++r) j.path[r] = take(ncell);
            j.disp = reinterpret_cast<float*>(take(npix * 4));
            j.W = p.width; j.H = p.height; j.D = p.disp_range; j.min_d = p.min_disparity;
            j.roi_x = std::max(0, p.min_disparity);
            j.is_right = v;
            j.cs_type = opt->cs_type;
            j.p1 = opt->p1; j.p2 = opt->p2_init;
            j.uniqueness = opt->unique_threshold;
            (v ? right_jobs : left_jobs).push_back(j);
            (v ? post.raw_r : post.raw_l) = j.disp;
        }
        if (!lr) {
            // the right census is still needed by the left cost: a census-only job
            SgmJob j = left_jobs.back();
            j.img = p.d_right; j.census = cs[1]; j.census_other = cs[0]; j.is_right = 1;
            right_jobs.push_back(j);
        }
        post.out_l = p.d_disp_left;
        post.out_r = p.d_disp_right;
        post.W = p.width; post.H = p.height;
        post.roi_x = std::max(0, p.min_disparity);
        post.do_lr = lr ? 1 : 0;
        post.do_median = opt->do_median_filter ? 1 : 0;
        post.lr_thr = opt->lr_threshold;
        posts.push_back(post);
        b->want_right = b->want_right || p.d_disp_right != nullptr;
    }
    std::vector<SgmJob> jobs = left_jobs;
    jobs.insert(jobs.end(), right_jobs.begin(), right_jobs.end());
    if (upload_plan(jobs, posts, &b->plan, &b->plan_block, &b->err)) {
        g_last_error = b->err;
        free_batch(b);
        return nullptr;
    }
    if (!lr) restrict_matching(&b->plan, jobs, n);   // only the left views are matched
    // the matched jobs: all of them with the LR check, else the n left views
    b->jobs.assign(jobs.begin(), jobs.begin() + (lr ? jobs.size() : (size_t)n));
    b->signature = batch_signature(b->device, *opt, pairs, n, b->want_right);
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess ||
        (b->plan.wide ? wide_occupancy(b->plan, &b->per_cu)
                      : persistent_occupancy(b->plan, &b->per_cu)) != hipSuccess) {
        fail(nullptr, SGM_E_HIP, "device query failed");
        free_batch(b);
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        auto it = g_sched_cache.find(b->signature);
        if (it != g_sched_cache.end() && apply_schedule(b, it->second.first, it->second.second)) {
            g_last_error = b->err;
            free_batch(b);
            return nullptr;
        }
    }
    return b;
}

int sgm_batch_run(sgm_batch* b, void* stream) {
    if (!b) return fail(nullptr, SGM_E_INVALID_ARG, "batch is NULL");
    HIP_TRY(&b->err, hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t ev[6] = {};
    if (b->use_graph) return run_graph(b, s);
    if (b->timing) {
        for (int k = 0; k < 6; ++k) HIP_TRY(&b->err, take_event(b, &ev[k]));
        b->marks.insert(b->marks.end(), ev, ev + 6);
    }
    HIP_TRY(&b->err, enqueue(b, s, b->timing ? ev : nullptr));
    return SGM_OK;
}

int sgm_batch_set_timing(sgm_batch* b, int32_t on) {
    if (!b) return fail(nullptr, SGM_E_INVALID_ARG, "batch is NULL");
    if (on && b->use_graph) return fail(&b->err, SGM_E_INVALID_ARG, "timing is off in graph mode");
    b->timing = on != 0;
    return SGM_OK;
}

int sgm_batch_use_graph(sgm_batch* b, int32_t on) {
    if (!b) return fail(nullptr, SGM_E_INVALID_ARG, "batch is NULL");
    if (on && b->timing) return fail(&b->err, SGM_E_INVALID_ARG, "timing is off in graph mode");
    b->use_graph = on != 0;
    return SGM_OK;
}

int sgm_batch_set_schedule(sgm_batch* b, int32_t policy, int32_t waves) {
    if (!b) return fail(nullptr, SGM_E_INVALID_ARG, "batch is NULL");
    return apply_schedule(b, policy, waves);
}

int sgm_batch_get_schedule(const sgm_batch* b, int32_t* policy, int32_t* waves) {
    if (!b || !policy || !waves) return fail(nullptr, SGM_E_INVALID_ARG, "NULL argument");
    *policy = b->policy;
    *waves = b->waves;
    return SGM_OK;
}

int sgm_batch_autotune(sgm_batch* b, void* stream, int32_t reps, int32_t force, int32_t* policy,
                       int32_t* waves, float* ms_per_run) {
    if (!b) return fail(nullptr, SGM_E_INVALID_ARG, "batch is NULL");
    if (reps <= 0) return fail(&b->err, SGM_E_INVALID_ARG, "reps must be positive");
    HIP_TRY(&b->err, hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto report = [&](float ms) {
        if (policy) *policy = b->policy;
        if (waves) *waves = b->waves;
        if (ms_per_run) *ms_per_run = ms;
        return SGM_OK;
    };
    if (!force) {
        std::pair<int, int> hit{-1, 0};
        {
            std::lock_guard<std::mutex> lk(g_sched_mu);
            auto it = g_sched_cache.find(b->signature);
            if (it != g_sched_cache.end()) hit = it->second;
        }
        if (hit.first >= 0) {
            const int rc = apply_schedule(b, hit.first, hit.second);
            return rc ? rc : report(-1.f);
        }
    }
    const int full = b->per_cu * b->n_cu;
    const std::vector<std::pair<int, int>> cand = {
        {SGM_SCHED_GRID, 0},
        {SGM_SCHED_LONGEST_FIRST, full},
        {SGM_SCHED_LONGEST_FIRST, std::max(1, full / 2)},
        {SGM_SCHED_LONGEST_FIRST, 2 * full},
        {SGM_SCHED_JOB_MAJOR, full},
    };
    const bool was_timing = b->timing, was_graph = b->use_graph;
    b->timing = false;
    b->use_graph = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(&b->err, hipEventCreate(&e0));
    HIP_TRY(&b->err, hipEventCreate(&e1));
    float best = 0.f;
    std::pair<int, int> pick = cand[0];
    int rc = SGM_OK;
    for (size_t c = 0; c < cand.size() && rc == SGM_OK; ++c) {
        if ((rc = apply_schedule(b, cand[c].first, cand[c].second)) != SGM_OK) break;
        hipError_t e = enqueue(b, s, nullptr);   // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        for (int r = 0; r < reps && e == hipSuccess; ++r) e = enqueue(b, s, nullptr);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = fail(&b->err, SGM_E_HIP, std::string("autotune run: ") + hipGetErrorString(e));
            break;
        }
        ms /= (float)reps;
        if (c == 0 || ms < best) {
            best = ms;
            pick = {cand[c].first, b->waves};
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->timing = was_timing;
    b->use_graph = was_graph;
    if (rc != SGM_OK) return rc;
    if ((rc = apply_schedule(b, pick.first, pick.second)) != SGM_OK) return rc;
    {
        std::lock_guard<std::mutex> lk(g_sched_mu);
        g_sched_cache[b->signature] = pick;
    }
    return report(best);
}

int sgm_batch_stage_times(sgm_batch* b, float* ms, int32_t* runs) {
    if (!b || !ms) return fail(nullptr, SGM_E_INVALID_ARG, "NULL argument");
    HIP_TRY(&b->err, hipSetDevice(b->device));
    for (int k = 0; k < 5; ++k) ms[k] = 0.f;
    const int nrun = (int)(b->marks.size() / 6);
    if (nrun > 0) HIP_TRY(&b->err, hipEventSynchronize(b->marks.back()));
    for (int r = 0; r < nrun; ++r)
        for (int k = 0; k < 5; ++k) {
            float t = 0.f;
            HIP_TRY(&b->err, hipEventElapsedTime(&t, b->marks[6 * r + k], b->marks[6 * r + k + 1]));
            ms[k] += t;
        }
    b->pool.insert(b->pool.end(), b->marks.begin(), b->marks.end());
    b->marks.clear();
    if (runs) *runs = nrun;
    return SGM_OK;
}

size_t sgm_batch_device_bytes(const sgm_batch* b) { return b ? b->arena_size : 0; }

void sgm_batch_destroy(sgm_batch* b) {
    if (b) free_batch(b);
}

}  // extern "C"

// ------------------------------------------------------------ stage entries
namespace {

struct TmpPlan {
    void* block = nullptr;
    StagePlan plan;
    ~TmpPlan() { if (block) (void)hipFree(block); }
};

// A view's region of interest (StereoROI_T x, y, w, h; NULL = the whole
// matched region) clipped to the image and to the columns with a cost.
struct Rect {
    int x0, y0, x1, y1;
    bool any() const { return x0 < x1 && y0 < y1; }
};

Rect roi_rect(const sgm_roi* r, int W, int H, int min_d) {
    const int lo = std::max(0, min_d);
    if (!r) return {lo, 0, W, H};
    return {std::max(r->x, lo), std::max(r->y, 0),
            (int)std::min<int64_t>((int64_t)r->x + r->w, W),
            (int)std::min<int64_t>((int64_t)r->y + r->h, H)};
}

// StereoCudaImpl::Match with regions of interest (StereoCudaImpl.cpp:418-446,
// 463-498; cusgm_ca.cu:3984-4000 and 4180-4189): the census and each view's
// cost cover the whole image; the view's rectangle is then packed into a
// cost volume of its own and its four directions and WTA run over that block
// only (each path starts on the rectangle's edge; the WTA's border rules use
// image coordinates through the job's frame), the rest of the raw map being
// INVALID; the LR check and the median cover the whole maps and peak removal
// the left rectangle's rows.  Uses the stored-cost kernels.
int run_roi(sgm_handle* h, const sgm_roi* rl, const sgm_roi* rr, bool want_right, hipStream_t s) {
    const bool lr = h->opt.do_lr_check != 0;
    const int W = h->W, H = h->H, D = h->D;
    HIP_TRY(&h->err, launch_census(h->plan, s));
    TmpPlan plans[4];
    Rect left_rect{0, 0, 0, 0};
    for (int v = 0; v < (lr ? 2 : 1); ++v) {
        const Rect r = roi_rect(v == 0 ? rl : rr, W, H, h->min_d);
        if (v == 0) left_rect = r;
        float* raw = h->view[v].raw;
        HIP_TRY(&h->err, launch_invalid_columns(raw, W, H, W, s));
        if (!r.any()) continue;
        // the view's cost over the whole image
        const SgmJob full = make_job(h, v);
        if (int rc = upload_plan({full}, {}, &plans[2 * v].plan, &plans[2 * v].block, &h->err)) return rc;
        HIP_TRY(&h->err, launch_cost(plans[2 * v].plan, s));
        // the rectangle, packed into the other view's cost buffer (free: the
        // right view's cost is computed after the left view is done)
        uint8_t* packed = h->view[1 - v].cost;
        const int wc = r.x1 - r.x0, hc = r.y1 - r.y0;
        HIP_TRY(&h->err, launch_crop_volume(h->view[v].cost, packed, W, D, r.x0, r.y0, wc, hc, s));
        SgmJob c = full;
        c.cost = packed;
        c.W = wc; c.H = hc; c.roi_x = 0;
        c.fx0 = r.x0; c.fy0 = r.y0; c.fW = W; c.fH = H;
        if (int rc = upload_plan({c}, {}, &plans[2 * v + 1].plan, &plans[2 * v + 1].block, &h->err))
            return rc;
        HIP_TRY(&h->err, launch_aggregate(plans[2 * v + 1].plan, s));
        HIP_TRY(&h->err, launch_wta(plans[2 * v + 1].plan, s));
    }
    HIP_TRY(&h->err, launch_post(h->plan, s));
    if (lr && want_right) HIP_TRY(&h->err, launch_right_invalidate(h->plan, s));
    if (h->opt.do_remove_peaks && left_rect.any())
        HIP_TRY(&h->err, launch_remove_peaks(h->view[0].out + (size_t)left_rect.y0 * W, h->peaks, W,
                                             left_rect.y1 - left_rect.y0, peak_strips(),
                                             h->opt.peaks_ratio_threshold, s));
    // the plans' device tables are freed on return
    HIP_TRY(&h->err, hipStreamSynchronize(s));
    return SGM_OK;
}

SgmJob stage_job(int32_t W, int32_t H, int32_t min_d, int32_t D) {
    SgmJob j{};
    j.W = W; j.H = H; j.D = D; j.min_d = min_d;
    j.roi_x = std::max(0, min_d);
    return j;
}

int stage_check(int32_t W, int32_t H, int32_t min_d, int32_t D) {
    if (W <= 0 || H <= 0) return fail(nullptr, SGM_E_INVALID_ARG, "image size must be positive");
    if (!valid_range(D)) return fail(nullptr, SGM_E_INVALID_ARG, "disp_range must be 64..512, 2^n");
    if (std::max(0, min_d) >= W) return fail(nullptr, SGM_E_INVALID_ARG, "bad min_disparity");
    return SGM_OK;
}

int run_stage(const SgmJob& j, const SgmPost* post, hipError_t (*fn)(const StagePlan&, hipStream_t),
              void* stream) {
    TmpPlan t;
    std::vector<SgmJob> jobs{j};
    std::vector<SgmPost> posts;
    if (post) posts.push_back(*post);
    if (int rc = upload_plan(jobs, posts, &t.plan, &t.block, nullptr)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(nullptr, fn(t.plan, s));
    HIP_TRY(nullptr, hipStreamSynchronize(s));
    return SGM_OK;
}

}  // namespace

extern "C" {

int sgm_stage_census(const uint8_t* d_img, int32_t width, int32_t height, int32_t cs_type,
                     int32_t is_right, uint64_t* d_census, void* stream) {
    if (int rc = stage_check(width, height, 0, 64)) return rc;
    if (!d_img || !d_census) return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    SgmJob j = stage_job(width, height, 0, 64);
    j.img = d_img; j.census = d_census; j.cs_type = cs_type ? 1 : 0; j.is_right = is_right ? 1 : 0;
    return run_stage(j, nullptr, launch_census, stream);
}

int sgm_stage_cost(const uint64_t* d_census_base, const uint64_t* d_census_match,
                   int32_t width, int32_t height, int32_t min_disparity, int32_t disp_range,
                   int32_t left_to_right, int32_t cs_type, uint8_t* d_cost, void* stream) {
    if (int rc = stage_check(width, height, min_disparity, disp_range)) return rc;
    if (!d_census_base || !d_census_match || !d_cost)
        return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    SgmJob j = stage_job(width, height, min_disparity, disp_range);
    j.census = const_cast<uint64_t*>(d_census_base);
    j.census_other = d_census_match;
    j.cost = d_cost;
    j.is_right = left_to_right ? 0 : 1;
    j.cs_type = cs_type ? 1 : 0;
    return run_stage(j, nullptr, launch_cost, stream);
}

int sgm_stage_aggregate(const uint8_t* d_cost, int32_t width, int32_t height,
                        int32_t min_disparity, int32_t disp_range, int32_t p1, int32_t p2,
                        uint8_t* const d_paths[4], void* stream) {
    if (int rc = stage_check(width, height, min_disparity, disp_range)) return rc;
    if (!d_cost || !d_paths) return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    if (p1 < 0 || p2 < 0 || (int64_t)p1 + p2 > 65000)
        return fail(nullptr, SGM_E_INVALID_ARG, "P1 and P2 must be >= 0 with P1 + P2 <= 65000");
    SgmJob j = stage_job(width, height, min_disparity, disp_range);
    j.cost = const_cast<uint8_t*>(d_cost);
    for (int r = 0; r < 4; ++r) {
        if (!d_paths[r]) return fail(nullptr, SGM_E_INVALID_ARG, "NULL path pointer");
        j.path[r] = d_paths[r];
    }
    j.p1 = p1; j.p2 = p2;
    return run_stage(j, nullptr, launch_aggregate, stream);
}

int sgm_stage_aggregate_wta_otf(const uint64_t* d_census_base, const uint64_t* d_census_match,
                                int32_t width, int32_t height, int32_t min_disparity,
                                int32_t disp_range, int32_t is_right, int32_t cs_type, int32_t p1,
                                int32_t p2, float uniqueness, uint8_t* const d_paths[4],
                                float* d_disp, void* stream) {
    if (int rc = stage_check(width, height, min_disparity, disp_range)) return rc;
    if (!d_census_base || !d_census_match || !d_paths || !d_disp)
        return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    if (p1 < 0 || p2 < 0 || (int64_t)p1 + p2 > 65000)
        return fail(nullptr, SGM_E_INVALID_ARG, "P1 and P2 must be >= 0 with P1 + P2 <= 65000");
    SgmJob j = stage_job(width, height, min_disparity, disp_range);
    j.census = const_cast<uint64_t*>(d_census_base);
    j.census_other = d_census_match;
    j.is_right = is_right ? 1 : 0;
    j.cs_type = cs_type ? 1 : 0;
    if (!wide_ok(j))
        return fail(nullptr, SGM_E_UNSUPPORTED,
                    "the on-the-fly cost takes D <= 256 and not the 5x5 right view");
    for (int r : {(int)kUD, (int)kLR, (int)kRL}) {
        if (!d_paths[r]) return fail(nullptr, SGM_E_INVALID_ARG, "NULL path pointer");
        j.path[r] = d_paths[r];
    }
    j.disp = d_disp;
    j.p1 = p1; j.p2 = p2;
    j.uniqueness = uniqueness;
    // the fused WTA writes the columns from roi_x on; the ones left of them
    // are INVALID, as the WTA stage writes them
    HIP_TRY(nullptr, launch_invalid_columns(d_disp, width, height, j.roi_x,
                                            static_cast<hipStream_t>(stream)));
    return run_stage(j, nullptr, [](const StagePlan& p, hipStream_t s) {
        const hipError_t e = launch_wide3(p, s);
        return e != hipSuccess ? e : launch_wide_wta(p, s);
    }, stream);
}

int sgm_stage_wta(uint8_t* const d_paths[4], int32_t width, int32_t height,
                  int32_t min_disparity, int32_t disp_range, float uniqueness,
                  float* d_disp, void* stream) {
    if (int rc = stage_check(width, height, min_disparity, disp_range)) return rc;
    if (!d_paths || !d_disp) return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    SgmJob j = stage_job(width, height, min_disparity, disp_range);
    for (int r = 0; r < 4; ++r) {
        if (!d_paths[r]) return fail(nullptr, SGM_E_INVALID_ARG, "NULL path pointer");
        j.path[r] = d_paths[r];
    }
    j.disp = d_disp;
    j.uniqueness = uniqueness;
    return run_stage(j, nullptr, launch_wta, stream);
}

int sgm_stage_post(const float* d_raw_left, const float* d_raw_right, int32_t width,
                   int32_t height, int32_t do_lr_check, float lr_threshold,
                   int32_t do_median, float* d_out_left, float* d_out_right, void* stream) {
    if (width <= 0 || height <= 0) return fail(nullptr, SGM_E_INVALID_ARG, "bad size");
    if (!d_raw_left || !d_out_left || (do_lr_check && !d_raw_right))
        return fail(nullptr, SGM_E_INVALID_ARG, "NULL pointer");
    SgmPost p{};
    p.raw_l = d_raw_left; p.raw_r = d_raw_right; p.out_l = d_out_left; p.out_r = d_out_right;
    p.W = width; p.H = height; p.do_lr = do_lr_check ? 1 : 0; p.do_median = do_median ? 1 : 0;
    p.lr_thr = lr_threshold;
    SgmJob dummy = stage_job(width, height, 0, 64);
    TmpPlan t;
    std::vector<SgmJob> jobs{dummy};
    std::vector<SgmPost> posts{p};
    if (int rc = upload_plan(jobs, posts, &t.plan, &t.block, nullptr)) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(nullptr, launch_post(t.plan, s));
    if (d_out_right && do_lr_check) HIP_TRY(nullptr, launch_right_invalidate(t.plan, s));
    HIP_TRY(nullptr, hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_stage_remove_peaks(float* d_disp, int32_t width, int32_t height, int32_t strips,
                           float ratio, int32_t* d_scratch, void* stream) {
    if (!d_disp || !d_scratch || width <= 0 || height <= 0 || strips < 1 ||
        (int64_t)width * height > INT32_MAX || !(ratio >= 0.0f && ratio <= 1.0f))
        return fail(nullptr, SGM_E_INVALID_ARG, "remove_peaks: bad arguments");
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(nullptr, launch_remove_peaks(d_disp, d_scratch, width, height, strips, ratio, s));
    HIP_TRY(nullptr, hipStreamSynchronize(s));
    return SGM_OK;
}

int sgm_stage_remove_peaks_noalloc(float* d_disp, int32_t width, int32_t height, int32_t strips,
                                   float ratio, void* stream) {
    if (!d_disp || width <= 0 || height <= 0 || strips < 1 || !(ratio >= 0.0f && ratio <= 1.0f))
        return fail(nullptr, SGM_E_INVALID_ARG, "remove_peaks_noalloc: bad arguments");
    if ((int64_t)width * height > kPeaksNoAllocMaxPixels)
        return fail(nullptr, SGM_E_INVALID_ARG,
                    "remove_peaks_noalloc: the map has more than 2097152 pixels (the fixed "
                    "scratch holds 2^21); use sgm_stage_remove_peaks with a workspace");
    int* work = nullptr;
    HIP_TRY(nullptr, peaks_noalloc_scratch(&work));
    // every launch starts by writing the parent and size entries it reads, so
    // what an earlier call left in the scratch does not matter
    HIP_TRY(nullptr, launch_remove_peaks(d_disp, work, width, height, strips, ratio,
                                         static_cast<hipStream_t>(stream)));
    return SGM_OK;
}

}  // extern "C"
