// StereoCuda drop-in class: a pimpl over the C ABI (sgm_capi.h).
// Mirrors ethan-li-coding/FasterStereoCuda SgmStereoCuda/StereoCuda.cpp:1-118
// (same method set, same bool-return error convention).
#include "StereoCuda.h"

#include <cstdio>

#include "sgm_capi.h"

namespace {

sgm_option to_c(const CuSGMOption& o) {
    sgm_option c{};
    c.p1 = o.p1;
    c.p2_init = o.p2_init;
    c.p2_max = o.p2_max;
    c.cs_type = static_cast<int32_t>(o.cs_type);
    c.num_paths = o.num_paths;
    c.do_median_filter = o.do_median_filter ? 1 : 0;
    c.do_lr_check = o.do_lr_check ? 1 : 0;
    c.lr_threshold = o.lr_threshold;
    c.unique_threshold = o.unique_threshold;
    c.using_constant_p2 = o.using_constant_p2 ? 1 : 0;
    c.do_remove_peaks = o.do_remove_peaks ? 1 : 0;
    c.peaks_ratio_threshold = o.peaks_ratio_threshold;
    c.post_filter_type = static_cast<int32_t>(o.post_filter_type);
    c.morphology_type = static_cast<int32_t>(o.morphology_type);
    return c;
}

inline sgm_handle* H(void* p) { return static_cast<sgm_handle*>(p); }

}  // namespace

StereoCuda::StereoCuda() : impl_(sgm_create()) {}

StereoCuda::~StereoCuda() {
    if (impl_) {
        sgm_destroy(H(impl_));
        impl_ = nullptr;
    }
}

bool StereoCuda::Init(sint32 width, sint32 height, sint32 min_disparity, sint32 disp_range,
                      CuSGMOption option, bool print_log) const {
    if (!impl_) return false;
    const sgm_option c = to_c(option);
    const int rc = sgm_init(H(impl_), width, height, min_disparity, disp_range, &c, print_log);
    if (rc != SGM_OK && print_log) std::printf("StereoCuda::Init: %s\n", sgm_last_error(H(impl_)));
    return rc == SGM_OK;
}

bool StereoCuda::Match(uint8* img_left, uint8* img_right, float32* disp_left, float32* disp_right,
                       sint16* init_disp_left, StereoROI_T* ste_roi_left,
                       StereoROI_T* ste_roi_right) const {
    if (!impl_) return false;
    return sgm_match(H(impl_), img_left, img_right, disp_left, disp_right, init_disp_left,
                     reinterpret_cast<const sgm_roi*>(ste_roi_left),
                     reinterpret_cast<const sgm_roi*>(ste_roi_right)) == SGM_OK;
}

bool StereoCuda::Init2(sint32 width, sint32 height, sint32 min_disparity, sint32 disp_range,
                       CuSGMOption option, CamParam_T cam_param, bool print_log) const {
    if (!impl_) return false;
    const sgm_option c = to_c(option);
    const sgm_camera cam{cam_param.x0_left,  cam_param.y0_left, cam_param.x0_right,
                         cam_param.y0_right, cam_param.focus,   cam_param.baseline};
    const int rc = sgm_init2(H(impl_), width, height, min_disparity, disp_range, &c, &cam, print_log);
    if (rc != SGM_OK && print_log) std::printf("StereoCuda::Init2: %s\n", sgm_last_error(H(impl_)));
    return rc == SGM_OK;
}

bool StereoCuda::Match2(uint8* img_left, uint8* img_right, float32* depth_left,
                        sint16* init_disp_left, StereoROI_T* ste_roi_left,
                        StereoROI_T* ste_roi_right) const {
    if (!impl_) return false;
    return sgm_match2(H(impl_), img_left, img_right, depth_left, init_disp_left,
                      reinterpret_cast<const sgm_roi*>(ste_roi_left),
                      reinterpret_cast<const sgm_roi*>(ste_roi_right)) == SGM_OK;
}

void StereoCuda::SetMinDisparity(const sint32& min_disparty) {
    if (impl_) sgm_set_min_disparity(H(impl_), min_disparty);
}

sint32 StereoCuda::GetDispartyRange() { return impl_ ? sgm_get_disparity_range(H(impl_)) : 0; }

void StereoCuda::Release() const {
    if (impl_) sgm_release(H(impl_));
}

const char* StereoCuda::LastError() const { return sgm_last_error(H(impl_)); }

void StereoCuda::GetRoiFromDispMap(float32* disp_ptr, sint32 width, sint32 height,
                                   StereoROI_T& ste_roi) {
    sgm_roi_from_disparity(disp_ptr, width, height, reinterpret_cast<sgm_roi*>(&ste_roi));
}

float32 StereoCuda::get_invad_float() { return sgm_invalid_float(); }
sint16 StereoCuda::get_invad_short() { return sgm_invalid_short(); }
sint16 StereoCuda::get_level_range() { return sgm_level_range(); }

bool StereoCuda::MallocPageLockedPtr(void** ptr, size_t size) {
    return sgm_host_alloc(ptr, size) == SGM_OK;
}

bool StereoCuda::FreePageLockedPtr(void* ptr) { return sgm_host_free(ptr) == SGM_OK; }
