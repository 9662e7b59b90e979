/*
 * FasterStereoCuda.h -- coarse-to-fine drop-in class (MI355X build).
 *
 * Same class, nested option structs (field names, order and defaults) and
 * member signatures as the reference's FasterStereoCuda/FasterStereoCuda.h:
 * 24-127, so code written against that header compiles unchanged.  The
 * implementation (libFasterStereoCuda.so) is a pimpl over the pyramid entry
 * points of sgm_capi.h; every level runs on the GPU.
 */
#ifndef FASTERSTEREOCUDA_STEREO_H_
#define FASTERSTEREOCUDA_STEREO_H_

#include <stddef.h>

#ifdef _WIN32
#ifndef FASTERSTEREOCUDA_EXPORTS
#define FASTERSTEREOCUDA_DLL __declspec(dllimport)
#else
#define FASTERSTEREOCUDA_DLL __declspec(dllexport)
#endif
#else
#define FASTERSTEREOCUDA_DLL __attribute__((visibility("default")))
#endif

class FASTERSTEREOCUDA_DLL FasterStereoCuda {
public:
    /* rectified-pair geometry (reference :28-37), read by the depth
     * interface (Init2 / Match2) */
    typedef struct EpipolarStereoParam {
        double x0_left, y0_left;
        double x0_right, y0_right;
        double baseline;
        double focus;
        EpipolarStereoParam() : x0_left(0), y0_left(0), x0_right(0), y0_right(0), baseline(0), focus(0) {}
    } epi_ste_t;

    /* disparity-map options (reference :39-51) */
    typedef struct StereoOption1 {
        int width;
        int height;
        int min_disp;
        int max_disp;
        int num_layers;
        bool do_lr_check;
        bool do_rm_peaks;    /* peak removal on the GPU after the coarsest level's
                                filters, with the reference's ratio per level
                                count; segments are whole-image connected
                                components (no per-thread strips) */
        bool do_smooth;      /* must be false: the Gauss post filter is not built
                                (Init returns false) */
        StereoOption1()
            : width(0), height(0), min_disp(0), max_disp(0), num_layers(1), do_lr_check(true),
              do_rm_peaks(true), do_smooth(false) {}
    } ste_opt1_t;

    /* depth-map options (reference :54-69) */
    typedef struct StereoOption2 {
        int width;
        int height;
        float min_depth;
        float max_depth;
        int num_layers;
        bool do_lr_check;
        bool do_rm_peaks;
        bool do_smooth;
        epi_ste_t epi;
        StereoOption2()
            : width(0), height(0), min_depth(0), max_depth(0), num_layers(1), do_lr_check(true),
              do_rm_peaks(true), do_smooth(false) {}
    } ste_opt2_t;

    FasterStereoCuda();
    ~FasterStereoCuda();
    FasterStereoCuda(const FasterStereoCuda&) = delete;
    FasterStereoCuda& operator=(const FasterStereoCuda&) = delete;

    static bool MallocPageLockedPtr(void** ptr, size_t size);
    static bool FreePageLockedPtr(void* ptr);

    /* Builds the level pyramid for option (reference :94, FasterStereoCuda.cpp:203-295):
     * 5x5 census, P1 = 8, P2 = 32, uniqueness 0.95 at level 0 and 1.0 above,
     * LR check at the coarsest level, 3x3 median, dilation of the coarsest
     * level, 32-disparity windows on the finer levels. */
    bool Init(const ste_opt1_t& option) const;
    /* Depth form (reference FasterStereoCuda.cpp:297-406): the disparity
     * interval comes from [min_depth, max_depth] and the camera in
     * option.epi; otherwise as Init. */
    bool Init2(const ste_opt2_t& option) const;
    void Release() const;

    /* Blocking match of a rectified uint8 pair (width x height, row-major)
     * into a float disparity map; invalid pixels hold -999999.0f. */
    bool Match(const unsigned char* bytes_left, const unsigned char* bytes_right,
               float* disparity_data) const;
    /* Match, then depth = baseline * focus / (disp + x0_right - x0_left) of
     * the final map on the GPU (float64 arithmetic); invalid pixels hold
     * -999999.0f. */
    bool Match2(const unsigned char* bytes_left, const unsigned char* bytes_right,
                float* depth_data) const;

    /* Message of the last failure on this object (not in the reference). */
    const char* LastError() const;

private:
    void* impl_;
};

#endif  // FASTERSTEREOCUDA_STEREO_H_
