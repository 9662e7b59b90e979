// Image files of the stereo console (see image_io.h): a PNG decoder and
// encoder over zlib, PGM and PFM.
#include "image_io.h"

#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace fsc {
namespace {

bool slurp(const std::string& path, std::string* out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    *out = ss.str();
    return true;
}

uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

void put_be32(std::string* s, uint32_t v) {
    for (int k = 3; k >= 0; --k) s->push_back((char)((v >> (8 * k)) & 0xFF));
}

bool fail(std::string* err, const std::string& m) {
    if (err) *err = m;
    return false;
}

// 15-bit fixed-point BT.601 luma, rounded (the factors sum to 2^15)
constexpr uint32_t kRed = 9797, kGreen = 19234, kBlue = 3737;
static_assert(kRed + kGreen + kBlue == 32768, "luma factors must sum to 2^15");

inline uint32_t luma(uint32_t r, uint32_t g, uint32_t b) {
    return (kRed * r + kGreen * g + kBlue * b + 16384) >> 15;
}

struct PngInfo {
    uint32_t w = 0, h = 0;
    int depth = 0, color = 0, interlace = 0;
    int channels = 0;                  // samples per pixel
    std::vector<uint8_t> palette;      // RGB triples
};

int channels_of(int color) {
    switch (color) {
        case 0: return 1;   // gray
        case 2: return 3;   // RGB
        case 3: return 1;   // palette index
        case 4: return 2;   // gray + alpha
        case 6: return 4;   // RGBA
        default: return 0;
    }
}

bool depth_ok(int color, int depth) {
    switch (color) {
        case 0: return depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16;
        case 3: return depth == 1 || depth == 2 || depth == 4 || depth == 8;
        case 2: case 4: case 6: return depth == 8 || depth == 16;
        default: return false;
    }
}

uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

// Reverses the filter of every row of one (sub-)image in place.  `data`
// holds rows of 1 + stride bytes (filter type, then the filtered bytes).
bool unfilter(uint8_t* data, size_t rows, size_t stride, size_t bpp, std::string* err) {
    std::vector<uint8_t> zero(stride, 0);
    const uint8_t* prev = zero.data();
    for (size_t y = 0; y < rows; ++y) {
        uint8_t* row = data + y * (stride + 1);
        const int type = row[0];
        uint8_t* cur = row + 1;
        switch (type) {
            case 0: break;
            case 1:
                for (size_t i = bpp; i < stride; ++i) cur[i] = (uint8_t)(cur[i] + cur[i - bpp]);
                break;
            case 2:
                for (size_t i = 0; i < stride; ++i) cur[i] = (uint8_t)(cur[i] + prev[i]);
                break;
            case 3:
                for (size_t i = 0; i < stride; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    cur[i] = (uint8_t)(cur[i] + ((left + prev[i]) >> 1));
                }
                break;
            case 4:
                for (size_t i = 0; i < stride; ++i) {
                    const int left = i >= bpp ? cur[i - bpp] : 0;
                    const int upleft = i >= bpp ? prev[i - bpp] : 0;
                    cur[i] = (uint8_t)(cur[i] + paeth(left, prev[i], upleft));
                }
                break;
            default:
                return fail(err, "PNG: unknown filter type " + std::to_string(type));
        }
        prev = cur;
    }
    return true;
}

// Sample k (0-based, in units of `depth` bits) of an unfiltered row.
inline uint32_t sample(const uint8_t* row, size_t k, int depth) {
    switch (depth) {
        case 16: return (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
        case 8: return row[k];
        default: {
            const size_t bit = k * depth;
            const int shift = 8 - depth - (int)(bit & 7);
            return (row[bit >> 3] >> shift) & ((1u << depth) - 1);
        }
    }
}

// Gray value of pixel x of an unfiltered row.
bool gray_of(const PngInfo& in, const uint8_t* row, size_t x, uint8_t* out, std::string* err) {
    const int d = in.depth;
    switch (in.color) {
        case 0:
        case 4: {
            const uint32_t v = sample(row, x * in.channels, d);
            // low depths scale to 8 bits by bit replication (exact multiples)
            *out = d == 16 ? (uint8_t)(v >> 8) : d == 8 ? (uint8_t)v : (uint8_t)(v * (255 / ((1u << d) - 1)));
            return true;
        }
        case 3: {
            const uint32_t idx = sample(row, x, d);
            if (3 * idx + 2 >= in.palette.size()) return fail(err, "PNG: palette index out of range");
            const uint8_t* c = &in.palette[3 * idx];
            *out = (uint8_t)luma(c[0], c[1], c[2]);
            return true;
        }
        case 2:
        case 6: {
            const size_t k = x * in.channels;
            const uint32_t r = sample(row, k, d), g = sample(row, k + 1, d), b = sample(row, k + 2, d);
            const uint32_t y = luma(r, g, b);
            *out = (uint8_t)(d == 16 ? (y >> 8) : y);
            return true;
        }
        default:
            return fail(err, "PNG: bad colour type");
    }
}

bool read_png(const std::string& file, std::vector<uint8_t>* px, int* w, int* h, std::string* err) {
    static const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    const uint8_t* p = reinterpret_cast<const uint8_t*>(file.data());
    const size_t n = file.size();
    if (n < 8 || std::memcmp(p, kSig, 8) != 0) return fail(err, "not a PNG file");
    PngInfo in;
    std::string idat;
    bool have_ihdr = false, have_iend = false;
    size_t pos = 8;
    while (pos + 12 <= n && !have_iend) {
        const uint32_t len = be32(p + pos);
        if (len > n - pos - 12) return fail(err, "PNG: truncated chunk");
        const uint8_t* type = p + pos + 4;
        const uint8_t* data = p + pos + 8;
        const uint32_t crc = be32(data + len);
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != crc)
            return fail(err, "PNG: chunk CRC mismatch");
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) return fail(err, "PNG: bad IHDR");
            in.w = be32(data);
            in.h = be32(data + 4);
            in.depth = data[8];
            in.color = data[9];
            in.interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || in.interlace > 1)
                return fail(err, "PNG: unknown compression, filter or interlace method");
            in.channels = channels_of(in.color);
            if (!in.channels || !depth_ok(in.color, in.depth))
                return fail(err, "PNG: bad colour type / bit depth");
            if (in.w == 0 || in.h == 0 || in.w > (1u << 16) || in.h > (1u << 16))
                return fail(err, "PNG: image size out of range");
            have_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            in.palette.assign(data, data + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.append(reinterpret_cast<const char*>(data), len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            have_iend = true;
        } else if (!(type[0] & 0x20)) {
            return fail(err, "PNG: unknown critical chunk");
        }
        pos += 12 + (size_t)len;
    }
    if (!have_ihdr || idat.empty()) return fail(err, "PNG: missing IHDR or IDAT");
    if (in.color == 3 && in.palette.empty()) return fail(err, "PNG: palette image without PLTE");

    // sub-images: one, or the seven Adam7 passes
    struct Pass { uint32_t x0, y0, dx, dy; };
    static const Pass kFull[1] = {{0, 0, 1, 1}};
    static const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                   {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass* passes = in.interlace ? kAdam7 : kFull;
    const int npass = in.interlace ? 7 : 1;
    const size_t bits = (size_t)in.channels * in.depth;
    const size_t bpp = std::max<size_t>(1, bits / 8);
    size_t total = 0;
    for (int k = 0; k < npass; ++k) {
        const size_t pw = (in.w + passes[k].dx - 1 - passes[k].x0) / passes[k].dx;
        const size_t ph = (in.h + passes[k].dy - 1 - passes[k].y0) / passes[k].dy;
        if (in.w > passes[k].x0 && in.h > passes[k].y0 && pw && ph) total += ph * (1 + (pw * bits + 7) / 8);
    }
    std::vector<uint8_t> raw(total);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) return fail(err, "zlib: inflateInit failed");
    zs.next_in = reinterpret_cast<Bytef*>(&idat[0]);
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    const int zr = inflate(&zs, Z_FINISH);
    const size_t produced = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || produced != raw.size())
        return fail(err, "PNG: image data is truncated or corrupt");

    px->assign((size_t)in.w * in.h, 0);
    size_t off = 0;
    for (int k = 0; k < npass; ++k) {
        const Pass& q = passes[k];
        if (in.w <= q.x0 || in.h <= q.y0) continue;
        const size_t pw = (in.w - q.x0 + q.dx - 1) / q.dx;
        const size_t ph = (in.h - q.y0 + q.dy - 1) / q.dy;
        if (!pw || !ph) continue;
        const size_t stride = (pw * bits + 7) / 8;
        uint8_t* data = raw.data() + off;
        if (!unfilter(data, ph, stride, bpp, err)) return false;
        for (size_t r = 0; r < ph; ++r) {
            const uint8_t* row = data + r * (stride + 1) + 1;
            uint8_t* dst = px->data() + (size_t)(q.y0 + r * q.dy) * in.w;
            for (size_t c = 0; c < pw; ++c)
                if (!gray_of(in, row, c, &dst[q.x0 + c * q.dx], err)) return false;
        }
        off += ph * (stride + 1);
    }
    *w = (int)in.w;
    *h = (int)in.h;
    return true;
}

// PGM header token (comments from '#' to the end of the line are skipped)
bool pgm_token(const std::string& s, size_t* pos, long* v) {
    while (*pos < s.size()) {
        if (std::isspace((unsigned char)s[*pos])) {
            ++*pos;
        } else if (s[*pos] == '#') {
            while (*pos < s.size() && s[*pos] != '\n') ++*pos;
        } else {
            break;
        }
    }
    const size_t b = *pos;
    while (*pos < s.size() && std::isdigit((unsigned char)s[*pos])) ++*pos;
    if (*pos == b || *pos - b > 9) return false;
    *v = std::stol(s.substr(b, *pos - b));
    return true;
}

bool read_pgm(const std::string& file, std::vector<uint8_t>* px, int* w, int* h, std::string* err) {
    const bool binary = file[1] == '5';
    size_t pos = 2;
    long W = 0, H = 0, maxv = 0;
    if (!pgm_token(file, &pos, &W) || !pgm_token(file, &pos, &H) || !pgm_token(file, &pos, &maxv) ||
        W <= 0 || H <= 0 || W > 65536 || H > 65536 || maxv <= 0 || maxv > 65535)
        return fail(err, "PGM: bad header");
    const size_t npx = (size_t)W * H;
    px->resize(npx);
    auto scale = [maxv](long v) { return (uint8_t)((v * 255 + maxv / 2) / maxv); };
    if (binary) {
        ++pos;   // the single blank after maxval
        const size_t bytes = maxv < 256 ? 1 : 2;
        if (file.size() < pos + npx * bytes) return fail(err, "PGM: truncated data");
        const uint8_t* d = reinterpret_cast<const uint8_t*>(file.data()) + pos;
        for (size_t i = 0; i < npx; ++i) {
            const long v = bytes == 1 ? d[i] : (long)d[2 * i] << 8 | d[2 * i + 1];
            if (v > maxv) return fail(err, "PGM: sample above maxval");
            (*px)[i] = maxv == 255 ? (uint8_t)v : scale(v);
        }
    } else {
        for (size_t i = 0; i < npx; ++i) {
            long v = 0;
            if (!pgm_token(file, &pos, &v) || v > maxv) return fail(err, "PGM: bad sample");
            (*px)[i] = maxv == 255 ? (uint8_t)v : scale(v);
        }
    }
    *w = (int)W;
    *h = (int)H;
    return true;
}

bool write_all(const std::string& path, const std::string& bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(bytes.data(), (std::streamsize)bytes.size());
    return bool(f);
}

void put_chunk(std::string* s, const char* type, const std::string& data) {
    put_be32(s, (uint32_t)data.size());
    std::string body(type, 4);
    body += data;
    s->append(body);
    put_be32(s, (uint32_t)crc32(crc32(0L, Z_NULL, 0),
                                reinterpret_cast<const Bytef*>(body.data()), (uInt)body.size()));
}

}  // namespace

bool read_gray(const std::string& path, std::vector<uint8_t>* px, int* w, int* h, std::string* err) {
    std::string file;
    if (!slurp(path, &file)) return fail(err, "cannot read " + path);
    if (file.size() >= 8 && static_cast<unsigned char>(file[0]) == 0x89 && file.compare(1, 3, "PNG") == 0)
        return read_png(file, px, w, h, err);
    if (file.size() >= 3 && file[0] == 'P' && (file[1] == '5' || file[1] == '2'))
        return read_pgm(file, px, w, h, err);
    return fail(err, path + ": not a PNG or PGM file");
}

bool write_png_gray(const std::string& path, const uint8_t* px, int w, int h) {
    std::string raw;
    raw.reserve((size_t)(w + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);   // filter: none
        raw.append(reinterpret_cast<const char*>(px) + (size_t)y * w, (size_t)w);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::string z(zlen, '\0');
    if (compress2(reinterpret_cast<Bytef*>(&z[0]), &zlen, reinterpret_cast<const Bytef*>(raw.data()),
                  (uLong)raw.size(), 6) != Z_OK)
        return false;
    z.resize(zlen);
    std::string out("\x89PNG\r\n\x1a\n", 8);
    std::string ihdr;
    put_be32(&ihdr, (uint32_t)w);
    put_be32(&ihdr, (uint32_t)h);
    ihdr += std::string("\x08\x00\x00\x00\x00", 5);   // 8-bit gray, deflate, no interlace
    put_chunk(&out, "IHDR", ihdr);
    put_chunk(&out, "IDAT", z);
    put_chunk(&out, "IEND", std::string());
    return write_all(path, out);
}

bool write_pgm(const std::string& path, const uint8_t* px, int w, int h) {
    std::string out = "P5\n" + std::to_string(w) + " " + std::to_string(h) + "\n255\n";
    out.append(reinterpret_cast<const char*>(px), (size_t)w * h);
    return write_all(path, out);
}

bool write_pfm(const std::string& path, const float* map, int w, int h) {
    std::string out = "Pf\n" + std::to_string(w) + " " + std::to_string(h) + "\n-1.0\n";
    static_assert(sizeof(float) == 4, "float must be 4 bytes");
    for (int y = h - 1; y >= 0; --y)   // bottom row first; -1.0 = little-endian
        out.append(reinterpret_cast<const char*>(map + (size_t)y * w), (size_t)w * 4);
    return write_all(path, out);
}

}  // namespace fsc
