// stereo_demo: drives the drop-in classes exactly as a caller of the
// reference libraries would (Init, Match, Release), on raw files.
//
//   stereo_demo LEFT.u8 RIGHT.u8 W H MIN_D RANGE OUT_LEFT.f32 [OUT_RIGHT.f32]
//               [--cs 0|1] [--p1 N] [--p2 N] [--uniq F] [--lr 0|1] [--lrthr F]
//               [--median 0|1] [--repeat N] [--layers N]
//               [--cam X0L,X0R,FOCUS,BASELINE] [--zrange MIN,MAX] [--peaks RATIO]
//               [--default-opt 1] [--init INIT.i16]
//
// Without --layers: StereoCuda (single level, the options above).  With
// --layers N: FasterStereoCuda (N-level pyramid over [MIN_D, MIN_D + RANGE),
// its own fixed per-level options; only --lr and --repeat apply).
// With --cam the output is depth instead of disparity, through Init2 /
// Match2: StereoCuda keeps MIN_D and RANGE; FasterStereoCuda takes its
// interval from --zrange (the depth interval) and ignores MIN_D and RANGE.
// --peaks turns on peak removal: StereoCuda with that ratio, FasterStereoCuda
// with do_rm_peaks (its own ratio; any RATIO value).
// --default-opt 1 starts StereoCuda from a default-constructed CuSGMOption
// (the other options then set single fields) instead of the benchmark's.
// --init passes an int16 W x H initial disparity map to StereoCuda::Match /
// Match2 (RANGE 32: the window search of the reference's pyramid levels).
// Inputs are tightly packed uint8 W x H; outputs tightly packed float32.
// Prints the average Match time over --repeat calls after the first.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "FasterStereoCuda.h"
#include "StereoCuda.h"

static bool read_file(const char* path, std::vector<uint8>& buf, size_t n) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    buf.resize(n);
    const size_t got = std::fread(buf.data(), 1, n, f);
    std::fclose(f);
    return got == n;
}

static bool write_file(const char* path, const void* data, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const size_t put = std::fwrite(data, 1, n, f);
    std::fclose(f);
    return put == n;
}

int main(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: %s LEFT RIGHT W H MIN_D RANGE OUT_LEFT [OUT_RIGHT] [options]\n",
                     argv[0]);
        return 2;
    }
    const int W = std::atoi(argv[3]), H = std::atoi(argv[4]);
    const int min_d = std::atoi(argv[5]), range = std::atoi(argv[6]);
    const char* out_left = argv[7];
    const char* out_right = nullptr;
    int a = 8;
    if (a < argc && std::strncmp(argv[a], "--", 2) != 0) out_right = argv[a++];

    CuSGMOption opt;
    bool defaults = false;
    for (int k = a; k + 1 < argc; k += 2)
        if (std::strcmp(argv[k], "--default-opt") == 0) defaults = std::atoi(argv[k + 1]) != 0;
    if (!defaults) {
        opt.cs_type = CuSGMOption::CS_9x7;
        opt.p1 = 10;
        opt.p2_init = 150;
        opt.unique_threshold = 0.95f;
        opt.do_lr_check = true;
        opt.lr_threshold = 1.0f;
        opt.do_median_filter = true;
        opt.using_constant_p2 = true;
    }
    const char* init_path = nullptr;
    int repeat = 1, layers = 0;
    bool depth = false;
    double cam[4] = {0, 0, 0, 0};      // x0_left, x0_right, focus, baseline
    float zrange[2] = {0.0f, 0.0f};
    for (; a + 1 < argc; a += 2) {
        const std::string k = argv[a];
        const char* v = argv[a + 1];
        if (k == "--cs") opt.cs_type = std::atoi(v) ? CuSGMOption::CS_9x7 : CuSGMOption::CS_5x5;
        else if (k == "--p1") opt.p1 = std::atoi(v);
        else if (k == "--p2") opt.p2_init = std::atoi(v);
        else if (k == "--uniq") opt.unique_threshold = (float)std::atof(v);
        else if (k == "--lr") opt.do_lr_check = std::atoi(v) != 0;
        else if (k == "--lrthr") opt.lr_threshold = (float)std::atof(v);
        else if (k == "--median") opt.do_median_filter = std::atoi(v) != 0;
        else if (k == "--repeat") repeat = std::atoi(v);
        else if (k == "--layers") layers = std::atoi(v);
        else if (k == "--default-opt") continue;
        else if (k == "--init") init_path = v;
        else if (k == "--peaks") {
            opt.do_remove_peaks = true;
            opt.peaks_ratio_threshold = (float)std::atof(v);
        }
        else if (k == "--cam") {
            depth = std::sscanf(v, "%lf,%lf,%lf,%lf", &cam[0], &cam[1], &cam[2], &cam[3]) == 4;
            if (!depth) { std::fprintf(stderr, "--cam needs X0L,X0R,FOCUS,BASELINE\n"); return 2; }
        } else if (k == "--zrange") {
            if (std::sscanf(v, "%f,%f", &zrange[0], &zrange[1]) != 2) {
                std::fprintf(stderr, "--zrange needs MIN,MAX\n");
                return 2;
            }
        }
        else { std::fprintf(stderr, "unknown option %s\n", k.c_str()); return 2; }
    }

    std::vector<uint8> left, right;
    const size_t npix = (size_t)W * H;
    if (!read_file(argv[1], left, npix) || !read_file(argv[2], right, npix)) {
        std::fprintf(stderr, "cannot read the input images\n");
        return 2;
    }
    std::vector<float32> dl(npix), dr(npix);
    std::vector<sint16> init;
    if (init_path) {
        std::vector<uint8> raw;
        if (!read_file(init_path, raw, npix * 2)) {
            std::fprintf(stderr, "cannot read the initial disparity map\n");
            return 2;
        }
        init.resize(npix);
        std::memcpy(init.data(), raw.data(), npix * 2);
    }
    sint16* const init_ptr = init_path ? init.data() : nullptr;

    if (layers > 0) {
        FasterStereoCuda fs;
        bool ok;
        if (depth) {
            FasterStereoCuda::ste_opt2_t o;
            o.width = W; o.height = H; o.min_depth = zrange[0]; o.max_depth = zrange[1];
            o.num_layers = layers; o.do_lr_check = opt.do_lr_check;
            o.do_rm_peaks = opt.do_remove_peaks; o.do_smooth = false;
            o.epi.x0_left = cam[0]; o.epi.x0_right = cam[1];
            o.epi.focus = cam[2]; o.epi.baseline = cam[3];
            ok = fs.Init2(o);
        } else {
            FasterStereoCuda::ste_opt1_t o;
            o.width = W; o.height = H; o.min_disp = min_d; o.max_disp = min_d + range;
            o.num_layers = layers; o.do_lr_check = opt.do_lr_check;
            o.do_rm_peaks = opt.do_remove_peaks; o.do_smooth = false;
            ok = fs.Init(o);
        }
        if (!ok) {
            std::fprintf(stderr, "Init failed: %s\n", fs.LastError());
            return 1;
        }
        double ms = 0.0;
        for (int r = 0; r < repeat; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
            if (!(depth ? fs.Match2(left.data(), right.data(), dl.data())
                        : fs.Match(left.data(), right.data(), dl.data()))) {
                std::fprintf(stderr, "Match failed: %s\n", fs.LastError());
                return 1;
            }
            const auto t1 = std::chrono::steady_clock::now();
            if (r > 0 || repeat == 1) ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        fs.Release();
        std::printf("match_ms %.4f\n", ms / (repeat > 1 ? repeat - 1 : 1));
        return write_file(out_left, dl.data(), npix * 4) ? 0 : 1;
    }

    StereoCuda stereo;
    CamParam_T cp{};
    cp.x0_left = cam[0]; cp.x0_right = cam[1]; cp.focus = cam[2]; cp.baseline = cam[3];
    if (!(depth ? stereo.Init2(W, H, min_d, range, opt, cp, false)
                : stereo.Init(W, H, min_d, range, opt, false))) {
        std::fprintf(stderr, "Init failed: %s\n", stereo.LastError());
        return 1;
    }
    double total_ms = 0.0;
    for (int r = 0; r < repeat; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!(depth ? stereo.Match2(left.data(), right.data(), dl.data(), init_ptr)
                    : stereo.Match(left.data(), right.data(), dl.data(),
                                   out_right ? dr.data() : nullptr, init_ptr))) {
            std::fprintf(stderr, "Match failed: %s\n", stereo.LastError());
            return 1;
        }
        const auto t1 = std::chrono::steady_clock::now();
        if (r > 0 || repeat == 1) total_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    stereo.Release();
    const int timed = repeat > 1 ? repeat - 1 : 1;
    std::printf("match_ms %.4f\n", total_ms / timed);
    if (!write_file(out_left, dl.data(), npix * 4)) return 1;
    if (out_right && !write_file(out_right, dr.data(), npix * 4)) return 1;
    return 0;
}
