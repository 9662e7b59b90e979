// FasterStereoCuda drop-in class: a pimpl over the pyramid entry points of
// the C ABI (sgm_capi.h).  Follows the reference's FasterStereoCuda.cpp:
// 203-406 (Init, Init2), 408-601 (Match, Match2), 693-766 (class shell) and
// 169-199 (per-level options): the level-dependent
// parts (uniqueness 1.0 above level 0, LR check only at the coarsest level,
// dilation of the coarsest level) are applied inside sgm_pyramid_*.
#include "FasterStereoCuda.h"

#include <algorithm>
#include <string>

#include "sgm_capi.h"

namespace {

struct Impl {
    sgm_pyramid* pyr = nullptr;
    std::string err;
};

inline Impl* I(void* p) { return static_cast<Impl*>(p); }

constexpr int kLevelRange = 32;   // window of the finer levels (reference LEVEL_RANGE)

sgm_option level_options(bool do_lr_check, bool do_rm_peaks, bool do_smooth, int num_layers) {
    sgm_option c{};
    c.p1 = 8;
    c.p2_init = 32;
    c.p2_max = 64;
    c.cs_type = 0;                       // 5x5 census
    c.num_paths = 4;
    c.do_median_filter = 1;
    c.do_lr_check = do_lr_check ? 1 : 0;
    c.lr_threshold = 1.0f;
    c.unique_threshold = 0.95f;
    c.using_constant_p2 = 1;
    c.do_remove_peaks = do_rm_peaks ? 1 : 0;
    // applied at the coarsest level only; its ratio by level count (:182-196)
    c.peaks_ratio_threshold = num_layers <= 1 ? 0.0005f : (num_layers <= 2 ? 0.001f : 0.005f);
    c.post_filter_type = do_smooth ? 1 : 0;
    c.morphology_type = 0;
    return c;
}

}  // namespace

FasterStereoCuda::FasterStereoCuda() : impl_(new Impl()) {}

FasterStereoCuda::~FasterStereoCuda() {
    Release();
    delete I(impl_);
    impl_ = nullptr;
}

bool FasterStereoCuda::MallocPageLockedPtr(void** ptr, size_t size) {
    return sgm_host_alloc(ptr, size) == SGM_OK;
}

bool FasterStereoCuda::FreePageLockedPtr(void* ptr) { return sgm_host_free(ptr) == SGM_OK; }

bool FasterStereoCuda::Init(const ste_opt1_t& option) const {
    Impl* m = I(impl_);
    Release();
    if (option.width <= 0 || option.height <= 0) {
        m->err = "image size error";
        return false;
    }
    const sgm_option c = level_options(option.do_lr_check, option.do_rm_peaks, option.do_smooth,
                                       std::max(1, option.num_layers));
    m->pyr = sgm_pyramid_create(option.width, option.height, option.min_disp, option.max_disp,
                                std::max(1, option.num_layers), kLevelRange, &c);
    if (!m->pyr) {
        m->err = sgm_last_error(nullptr);
        return false;
    }
    m->err.clear();
    return true;
}

bool FasterStereoCuda::Init2(const ste_opt2_t& option) const {
    Impl* m = I(impl_);
    Release();
    if (option.width <= 0 || option.height <= 0) {
        m->err = "image size error";
        return false;
    }
    const sgm_option c = level_options(option.do_lr_check, option.do_rm_peaks, option.do_smooth,
                                       std::max(1, option.num_layers));
    const sgm_camera cam{option.epi.x0_left,  option.epi.y0_left, option.epi.x0_right,
                         option.epi.y0_right, option.epi.focus,   option.epi.baseline};
    m->pyr = sgm_pyramid_create2(option.width, option.height, option.min_depth, option.max_depth,
                                 std::max(1, option.num_layers), kLevelRange, &c, &cam);
    if (!m->pyr) {
        m->err = sgm_last_error(nullptr);
        return false;
    }
    m->err.clear();
    return true;
}

void FasterStereoCuda::Release() const {
    Impl* m = I(impl_);
    if (m && m->pyr) {
        sgm_pyramid_destroy(m->pyr);
        m->pyr = nullptr;
    }
}

bool FasterStereoCuda::Match(const unsigned char* bytes_left, const unsigned char* bytes_right,
                             float* disparity_data) const {
    Impl* m = I(impl_);
    if (!m->pyr) {
        m->err = "not initialised";
        return false;
    }
    if (!bytes_left || !bytes_right || !disparity_data) {
        m->err = "NULL image or output pointer";
        return false;
    }
    if (sgm_pyramid_match(m->pyr, bytes_left, bytes_right, disparity_data) != SGM_OK) {
        m->err = sgm_last_error(nullptr);
        return false;
    }
    return true;
}

bool FasterStereoCuda::Match2(const unsigned char* bytes_left, const unsigned char* bytes_right,
                              float* depth_data) const {
    Impl* m = I(impl_);
    if (!m->pyr) {
        m->err = "not initialised";
        return false;
    }
    if (!bytes_left || !bytes_right || !depth_data) {
        m->err = "NULL image or output pointer";
        return false;
    }
    if (sgm_pyramid_match2(m->pyr, bytes_left, bytes_right, depth_data) != SGM_OK) {
        m->err = sgm_last_error(nullptr);
        return false;
    }
    return true;
}

const char* FasterStereoCuda::LastError() const { return I(impl_)->err.c_str(); }
