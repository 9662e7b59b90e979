// Image files of the stereo console.
//
// Reading gives the 8-bit grayscale bytes the matcher takes, row-major,
// tightly packed.  PNG (every colour type, bit depths 1-16, non-interlaced
// and Adam7) is decoded with zlib's inflate; binary and ASCII PGM are read
// directly.  Colour is reduced to gray with the ITU-R BT.601 weights 0.299 /
// 0.587 / 0.114 as 15-bit fixed-point factors, rounded to nearest:
//     gray = (9797 R + 19234 G + 3737 B + 16384) >> 15
// (the factors libpng derives for png_set_rgb_to_gray(0.299, 0.587), which
// the reference's OpenCV 3.2 imread(IMREAD_GRAYSCALE) requests; whether
// OpenCV's result equals this one on every file is not pinned).  Equal R, G
// and B give that value back.  16-bit samples keep their high byte; alpha is
// dropped; palette entries are expanded first.
//
// Writing: 8-bit gray PNG, and PFM (the float map, little-endian, rows stored
// bottom-up as the format defines).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fsc {

bool read_gray(const std::string& path, std::vector<uint8_t>* px, int* w, int* h, std::string* err);

bool write_png_gray(const std::string& path, const uint8_t* px, int w, int h);
bool write_pgm(const std::string& path, const uint8_t* px, int w, int h);
bool write_pfm(const std::string& path, const float* map, int w, int h);

}  // namespace fsc
