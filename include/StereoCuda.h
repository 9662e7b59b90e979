/*
 * StereoCuda.h -- single-level SGM drop-in class (MI355X build).
 *
 * Same class name and member signatures as the reference's
 * SgmStereoCuda/StereoCuda.h:14-100.  The implementation (libStereoCuda.so)
 * is a thin pimpl that calls only the C ABI in sgm_capi.h; HIP is never
 * visible to the caller.  Methods keep the reference's `const` qualifiers
 * (state lives behind the pimpl).
 */
#pragma once
#include "sgm_types.h"

#ifdef _WIN32
#ifdef STEREO_CUDA_EXPORT
#define STEREO_CUDA_DLL __declspec(dllexport)
#else
#define STEREO_CUDA_DLL __declspec(dllimport)
#endif
#else
#define STEREO_CUDA_DLL __attribute__((visibility("default")))
#endif

class STEREO_CUDA_DLL StereoCuda {
public:
    StereoCuda();
    ~StereoCuda();
    StereoCuda(const StereoCuda&) = delete;
    StereoCuda& operator=(const StereoCuda&) = delete;

    /* width x height pair, disparities [min_disparity, min_disparity + disp_range);
     * disp_range = 64 * 2^n (n = 0..3), or get_level_range() (32) for an
     * object that matches only around an initial disparity map. */
    bool Init(sint32 width, sint32 height, sint32 min_disparity, sint32 disp_range,
              CuSGMOption sgm_option, bool print_log = false) const;

    /* Blocking match of a rectified uint8 pair into float disparity maps
     * (INVALID_VALUE marks invalid pixels).  disp_right may be nullptr.
     * init_disp_left (sint16, INVALID = get_invad_short()) is taken by an
     * object initialised with disp_range 32, do_lr_check false and
     * using_constant_p2: it matches in a 32-disparity window starting at each
     * pixel's initial disparity, as the coarse-to-fine levels do.
     * ste_roi_left / ste_roi_right restrict each view's aggregation and WTA to
     * that rectangle (clipped to the image; paths start on its edges); pixels
     * outside it are INVALID.  A region together with init_disp_left is
     * refused (returns false). */
    bool Match(uint8* img_left, uint8* img_right, float32* disp_left,
               float32* disp_right = nullptr, sint16* init_disp_left = nullptr,
               StereoROI_T* ste_roi_left = nullptr, StereoROI_T* ste_roi_right = nullptr) const;

    /* Depth output (reference StereoCuda.h:57,69): Init as above plus the
     * camera; Match2 runs Match and converts the final left map on the GPU,
     * depth = baseline * focus / (disp + x0_right - x0_left) in float64,
     * with invalid pixels passed through as get_invad_float(). */
    bool Init2(sint32 width, sint32 height, sint32 min_disparity, sint32 disp_range,
               CuSGMOption sgm_option, CamParam_T cam_param, bool print_log = false) const;
    bool Match2(uint8* img_left, uint8* img_right, float32* depth_left,
                sint16* init_disp_left = nullptr, StereoROI_T* ste_roi_left = nullptr,
                StereoROI_T* ste_roi_right = nullptr) const;

    void SetMinDisparity(const sint32& min_disparty);
    sint32 GetDispartyRange();

    void Release() const;

    /* Message of the last failure on this object (not in the reference). */
    const char* LastError() const;

    static void GetRoiFromDispMap(float32* disp_ptr, sint32 width, sint32 height,
                                  StereoROI_T& ste_roi);
    static float32 get_invad_float();
    static sint16 get_invad_short();
    static sint16 get_level_range();
    static bool MallocPageLockedPtr(void** ptr, size_t size);
    static bool FreePageLockedPtr(void* ptr);

private:
    void* impl_;
};
