// stereo_console: the command-line front end of the FasterStereoCuda class,
// with the argument modes of the reference's FasterStereoConsole
// (FasterStereoConsole.cpp:80-337) and its option.xml files (option_xml.h).
//
//   stereo_console LEFT RIGHT OPTION.xml [flags]          one pair
//   stereo_console DIR_LEFT DIR_RIGHT EXT OPTION.xml [flags]
//                                                         every *.EXT file of the two
//                                                         folders, paired in name order
//   stereo_console opt                                    writes option.xml and option2.xml
//                                                         (default values) where missing
//   stereo_console --to-gray IMAGE OUT.pgm                the gray bytes an image is matched as
//   stereo_console --show-option OPTION.xml               the option values as read
//
// The root element of OPTION.xml chooses the form: <Option1> gives a
// disparity map through Init / Match, <Option2> a depth map through Init2 /
// Match2.  Images are PNG or PGM, reduced to 8-bit gray as image_io.h states.
// When the option file gives no size (0), the first left image's size is
// used; every image must have the option's size.
//
// Flags (the reference shows its results in windows instead):
//   --out DIR    writes DIR/<left name>_disp.pfm (the float map, invalid =
//                -999999) and DIR/<left name>_disp.png (8 bits: (v - min) /
//                (max - min) * 255 over the valid values, 0 where invalid or
//                where the left image is 0, as the reference's view)
//   --gray DIR   writes the gray bytes matched, DIR/<name>.pgm
//   --loops N    passes over the frames (default 1; the reference repeats
//                them until it is stopped)
//
// Prints, per frame, "frame i, timing: T ms fps = F" as the reference does
// (fps over the frames after the first).  Exits 0 on success, 1 on an error,
// 2 on bad arguments.
#include <dirent.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "FasterStereoCuda.h"
#include "image_io.h"
#include "option_xml.h"

namespace {

struct Frame {
    std::string left, right;
};

std::string base_name(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    std::string b = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = b.find_last_of('.');
    return dot == std::string::npos || dot == 0 ? b : b.substr(0, dot);
}

// the regular files "*.ext" of a folder, in byte order of their names
std::vector<std::string> list_folder(const std::string& dir, const std::string& ext) {
    std::vector<std::string> names;
    DIR* d = opendir(dir.c_str());
    if (!d) return names;
    const std::string suffix = "." + ext;
    while (const dirent* e = readdir(d)) {
        const std::string n = e->d_name;
        if (n.size() <= suffix.size() || n.compare(n.size() - suffix.size(), suffix.size(), suffix) != 0)
            continue;
        const std::string full = dir + (dir.back() == '/' ? "" : "/") + n;
        struct stat st;
        if (::stat(full.c_str(), &st) == 0 && S_ISREG(st.st_mode)) names.push_back(full);
    }
    closedir(d);
    std::sort(names.begin(), names.end());
    return names;
}

int usage(const char* prog) {
    std::printf("insufficient parameters!\n"
                "usage: %s LEFT RIGHT OPTION.xml [--out DIR] [--gray DIR] [--loops N]\n"
                "       %s DIR_LEFT DIR_RIGHT EXT OPTION.xml [--out DIR] [--gray DIR] [--loops N]\n"
                "       %s opt\n"
                "       %s --to-gray IMAGE OUT.pgm\n"
                "       %s --show-option OPTION.xml\n",
                prog, prog, prog, prog, prog);
    return 2;
}

// the reference's 8-bit view of a map (FasterStereoConsole.cpp:276-300)
std::vector<uint8_t> view8(const float* m, const uint8_t* left, int w, int h) {
    const size_t n = (size_t)w * h;
    float lo = 999999.0f, hi = -999999.0f;
    for (size_t i = 0; i < n; ++i)
        if (m[i] != -999999.0f) {
            lo = std::min(lo, m[i]);
            hi = std::max(hi, m[i]);
        }
    std::vector<uint8_t> out(n, 0);
    if (!(hi > lo)) return out;   // no valid value, or a single one: all 0
    for (size_t i = 0; i < n; ++i)
        if (m[i] != -999999.0f && left[i] != 0)
            out[i] = static_cast<uint8_t>((m[i] - lo) / (hi - lo) * 255);
    return out;
}

}  // namespace

int show_option(const std::string& path) {
    switch (fsc::option_kind(path)) {
        case fsc::OptionKind::kDisparity: {
            FasterStereoCuda::StereoOption1 o;
            if (!fsc::load_option(path, o)) break;
            std::printf("form=Option1\nwidth=%d\nheight=%d\nmin_disp=%d\nmax_disp=%d\nnum_layers=%d\n"
                        "do_lr_check=%d\ndo_rm_peaks=%d\ndo_smooth=%d\n",
                        o.width, o.height, o.min_disp, o.max_disp, o.num_layers, o.do_lr_check,
                        o.do_rm_peaks, o.do_smooth);
            return 0;
        }
        case fsc::OptionKind::kDepth: {
            FasterStereoCuda::StereoOption2 o;
            if (!fsc::load_option(path, o)) break;
            std::printf("form=Option2\nwidth=%d\nheight=%d\nmin_depth=%.9g\nmax_depth=%.9g\n"
                        "num_layers=%d\ndo_lr_check=%d\ndo_rm_peaks=%d\ndo_smooth=%d\n"
                        "x0_left=%.17g\ny0_left=%.17g\nx0_right=%.17g\ny0_right=%.17g\n"
                        "baseline=%.17g\nfocus=%.17g\n",
                        o.width, o.height, o.min_depth, o.max_depth, o.num_layers, o.do_lr_check,
                        o.do_rm_peaks, o.do_smooth, o.epi.x0_left, o.epi.y0_left, o.epi.x0_right,
                        o.epi.y0_right, o.epi.baseline, o.epi.focus);
            return 0;
        }
        default:
            break;
    }
    std::printf("loading parameters of algorithm failed!\n");
    return 1;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::strcmp(argv[1], "--to-gray") == 0) {
        std::vector<uint8_t> px;
        int w = 0, h = 0;
        std::string err;
        if (!fsc::read_gray(argv[2], &px, &w, &h, &err)) {
            std::printf("%s\n", err.c_str());
            return 1;
        }
        return fsc::write_pgm(argv[3], px.data(), w, h) ? 0 : 1;
    }
    if (argc == 3 && std::strcmp(argv[1], "--show-option") == 0) return show_option(argv[2]);
    std::vector<std::string> pos;
    std::string out_dir, gray_dir;
    int loops = 1;
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        if (a == "--out" || a == "--gray" || a == "--loops") {
            if (k + 1 >= argc) return usage(argv[0]);
            const std::string v = argv[++k];
            if (a == "--out") out_dir = v;
            else if (a == "--gray") gray_dir = v;
            else loops = std::atoi(v.c_str());
        } else if (a.compare(0, 2, "--") == 0) {
            std::printf("unknown flag %s\n", a.c_str());
            return usage(argv[0]);
        } else {
            pos.push_back(a);
        }
    }
    if (loops < 1) return usage(argv[0]);

    if (pos.size() == 1 && pos[0] == "opt") {
        // mode 3: option files with default values (FasterStereoConsole.cpp:113-121)
        FasterStereoCuda::StereoOption1 o1;
        FasterStereoCuda::StereoOption2 o2;
        fsc::load_option("option.xml", o1);
        fsc::load_option("option2.xml", o2);
        std::printf("initialize parameter file of algorithm done!\n");
        return 0;
    }
    if (pos.size() != 3 && pos.size() != 4) {
        if (pos.size() < 3) return usage(argv[0]);
        std::printf("parameter type error!\n");
        return 2;
    }
    const std::string opt_path = pos.back();

    // options (FasterStereoConsole.cpp:129-157)
    FasterStereoCuda::StereoOption1 opt1;
    FasterStereoCuda::StereoOption2 opt2;
    const fsc::OptionKind kind = fsc::option_kind(opt_path);
    if (kind == fsc::OptionKind::kNone ||
        (kind == fsc::OptionKind::kDisparity && !fsc::load_option(opt_path, opt1)) ||
        (kind == fsc::OptionKind::kDepth && !fsc::load_option(opt_path, opt2))) {
        std::printf("loading parameters of algorithm failed!\n");
        return 1;
    }
    const bool depth = kind == fsc::OptionKind::kDepth;

    // frames (FasterStereoConsole.cpp:159-173)
    std::vector<Frame> frames;
    if (pos.size() == 3) {
        frames.push_back({pos[0], pos[1]});
    } else {
        const std::vector<std::string> ls = list_folder(pos[0], pos[2]);
        const std::vector<std::string> rs = list_folder(pos[1], pos[2]);
        if (ls.size() != rs.size()) {
            std::printf("data error! %zu left and %zu right images\n", ls.size(), rs.size());
            return 1;
        }
        for (size_t i = 0; i < ls.size(); ++i) frames.push_back({ls[i], rs[i]});
    }
    if (frames.empty()) {
        std::printf("data error! no images\n");
        return 1;
    }
    std::vector<std::vector<uint8_t>> lefts(frames.size()), rights(frames.size());
    int W = 0, H = 0;
    for (size_t i = 0; i < frames.size(); ++i) {
        int wl = 0, hl = 0, wr = 0, hr = 0;
        std::string err;
        if (!fsc::read_gray(frames[i].left, &lefts[i], &wl, &hl, &err) ||
            !fsc::read_gray(frames[i].right, &rights[i], &wr, &hr, &err)) {
            std::printf("data error! %s\n", err.c_str());
            return 1;
        }
        if (i == 0) {
            W = wl;
            H = hl;
        }
        if (wl != W || hl != H || wr != W || hr != H) {
            std::printf("data error! %s / %s: %dx%d and %dx%d, expected %dx%d\n",
                        frames[i].left.c_str(), frames[i].right.c_str(), wl, hl, wr, hr, W, H);
            return 1;
        }
        if (!gray_dir.empty() &&
            (!fsc::write_pgm(gray_dir + "/" + base_name(frames[i].left) + ".pgm", lefts[i].data(), W, H) ||
             !fsc::write_pgm(gray_dir + "/" + base_name(frames[i].right) + ".pgm", rights[i].data(), W, H))) {
            std::printf("cannot write to %s\n", gray_dir.c_str());
            return 1;
        }
    }
    int& ow = depth ? opt2.width : opt1.width;
    int& oh = depth ? opt2.height : opt1.height;
    if (ow <= 0 && oh <= 0) {
        ow = W;
        oh = H;
    }
    if (ow != W || oh != H) {
        std::printf("data error! the images are %dx%d, the option file says %dx%d\n", W, H, ow, oh);
        return 1;
    }

    // algorithm (FasterStereoConsole.cpp:175-192)
    FasterStereoCuda stereo;
    if (!(depth ? stereo.Init2(opt2) : stereo.Init(opt1))) {
        std::printf("initialize algorithm failed: %s\n", stereo.LastError());
        return 1;
    }
    std::printf("initialize algorithm done!\n");

    // page-locked buffers, as the reference allocates them (:196-205)
    const size_t npx = (size_t)W * H;
    unsigned char* img_left = nullptr;
    unsigned char* img_right = nullptr;
    float* d_data = nullptr;
    if (!FasterStereoCuda::MallocPageLockedPtr(reinterpret_cast<void**>(&img_left), npx) ||
        !FasterStereoCuda::MallocPageLockedPtr(reinterpret_cast<void**>(&img_right), npx) ||
        !FasterStereoCuda::MallocPageLockedPtr(reinterpret_cast<void**>(&d_data), npx * sizeof(float))) {
        std::printf("page-locked allocation failed\n");
        return 1;
    }

    // matching (:221-326)
    int rc = 0;
    int count = 0;
    double all = 0.0, first = 0.0;
    for (int pass = 0; pass < loops && rc == 0; ++pass) {
        for (size_t i = 0; i < frames.size(); ++i) {
            std::memcpy(img_left, lefts[i].data(), npx);
            std::memcpy(img_right, rights[i].data(), npx);
            const auto t0 = std::chrono::steady_clock::now();
            const bool ok = depth ? stereo.Match2(img_left, img_right, d_data)
                                  : stereo.Match(img_left, img_right, d_data);
            const auto t1 = std::chrono::steady_clock::now();
            if (!ok) {
                std::printf("frame %zu: matching failed: %s\n", i, stereo.LastError());
                rc = 1;
                break;
            }
            const double ms =
                std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count() / 1000.0;
            all += ms;
            if (count == 0) first = ms;
            ++count;
            const double avg = count > 1 ? (all - first) / (count - 1) : ms;
            const double fps = int(1000.0 / avg * 100 + 0.5) / 100.0;
            std::printf("frame %zu, timing: %6.4lf ms fps = %6.4lf\n", i, ms, fps);
            if (!out_dir.empty() && pass == 0) {
                const std::string stem = out_dir + "/" + base_name(frames[i].left) + "_disp";
                const std::vector<uint8_t> v = view8(d_data, img_left, W, H);
                if (!fsc::write_pfm(stem + ".pfm", d_data, W, H) ||
                    !fsc::write_png_gray(stem + ".png", v.data(), W, H)) {
                    std::printf("cannot write to %s\n", out_dir.c_str());
                    rc = 1;
                    break;
                }
            }
        }
    }

    stereo.Release();
    FasterStereoCuda::FreePageLockedPtr(img_left);
    FasterStereoCuda::FreePageLockedPtr(img_right);
    FasterStereoCuda::FreePageLockedPtr(d_data);
    return rc;
}
