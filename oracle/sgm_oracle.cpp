/*
 * CPU oracle for the SGM stereo path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg load it.
 *
 * Restates the integer arithmetic of ethan-li-coding/FasterStereoCuda
 * (paths below are relative to that repository).  Built with
 * -O3 -fopenmp -ffp-contract=off and no fast-math so that the float steps
 * (uniqueness test, parabola subpixel, LR average) round exactly as IEEE
 * single precision does on the GPU.
 *
 * Decisions for reference behaviour that is racy or undefined:
 *   A1  WTA argmin tie -> the smallest d attaining the minimum
 *   A2  LR check reads the right map as it was before the check
 *   A3  median is out of place; out-of-image neighbours are INVALID;
 *       x = 0 and y = 0 pass through unchanged
 *   A5  WTA rows/columns the reference grid never covers -> INVALID
 *   A7  LR check runs over xl in [0, W)
 *   A8  the first pixel of each path is stored (L = C)
 *   A9  median filters every row
 *   A10 64-bit offsets
 *   A12 INVALID_COST=127 boundary terms without +P1, uint8 storage of the
 *       per-path costs, c3 = D-1 at d* = D-1 (D in {64,128}), +0.5 bias: kept
 *   A13 right-border census constant 0x00000000FFFFFFFF: kept
 *   A16 5x5 right-to-left cost skips the last pixel of every D-pixel block
 *       (cusgm_cc.cu:288); that pixel keeps the left-to-right cost the same
 *       buffer held: kept
 *   A17 with the LR check off the right map is INVALID everywhere
 *   A6  pyramid P2 divisor abs(b' - b + 1) is clamped to >= 1
 *   A11 vertical pyramid paths seed the previous base disparity from row 0
 *       for both directions (cusgm_ca.cu:1291): kept
 *   A15 the pyramid window (LEVEL_RANGE) is a parameter; CostAdjust uses it
 *   A18 init propagation with an odd level width clamps the coarse column
 *       (the reference reads one element past the coarse row)
 */
#include "sgm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline size_t idx2(int x, int y, int W) { return (size_t)y * (size_t)W + (size_t)x; }
inline size_t idx3(int x, int y, int d, int W, int D) {
    return ((size_t)y * (size_t)W + (size_t)x) * (size_t)D + (size_t)d;
}

inline uint32_t popcount64(uint64_t v) { return (uint32_t)__builtin_popcountll(v); }

}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Census transform.  cusgm_cc.cu:30-67 (5x5, radius 2x2) and :68-106
 * (9x7, radius 4x3).  Bits are shifted in MSB-first over the raster order
 * i = -rh..rh, j = -rw..rw, the centre included; a bit is 1 iff the
 * neighbour is strictly darker than the centre.  Border pixels take
 * invalid_val: 0 for the left view, 0xFFFFFFFF for the right view
 * (cusgm_cc.cu:450-464; for 9x7 the uint32 constant widens to
 * 0x00000000FFFFFFFF). */
void orc_census(const uint8_t* img, int W, int H, int cs_type, int is_right,
                uint64_t* out) {
    const int rw = cs_type ? 4 : 2;
    const int rh = cs_type ? 3 : 2;
    const uint64_t invalid = is_right ? 0xFFFFFFFFull : 0ull;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            uint64_t v = 0;
            if (x < rw || x >= W - rw || y < rh || y >= H - rh) {
                v = invalid;
            } else {
                const uint8_t c = img[idx2(x, y, W)];
                for (int i = -rh; i <= rh; ++i)
                    for (int j = -rw; j <= rw; ++j)
                        v = (v << 1) | (uint64_t)(img[idx2(x + j, y + i, W)] < c);
            }
            out[idx2(x, y, W)] = v;
        }
    }
}

/* Hamming cost volume without initial disparity.  cusgm_cc.cu:118-172 (u64)
 * and :238-293 (u32), launch geometry :472-523.
 *   left_to_right:  C[y][x][t] = popc(cb[y][x] ^ cm[y][x - min_d - t])
 *   right_to_left:  C[y][x][t] = popc(cb[y][x] ^ cm[y][x + min_d + t])
 * with cb = census of the view being matched and cm = the other view; a
 * partner column outside [0, W) is replaced by cm[y][0].  Only columns
 * x >= roi_x = max(0, min_d) are written (cusgm_cc.cu:486-487).
 * A16: for the 5x5 census the right-to-left kernel's loop stops one pixel
 * short of each D-pixel block (cusgm_cc.cu:288); such a pixel keeps what
 * the buffer held, which in StereoCudaImpl::Match is the left-to-right cost
 * of the same frame (cusgm_ca.cu:4191 reuses the buffer).  Pass that volume
 * as l2r_cost. */
void orc_cost(const uint64_t* cs_base, const uint64_t* cs_match, int W, int H,
              int min_d, int D, int left_to_right, int cs_type,
              const uint8_t* l2r_cost, uint8_t* cost) {
    const int roi_x = std::max(0, min_d);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        const uint64_t* cb = cs_base + (size_t)y * W;
        const uint64_t* cm = cs_match + (size_t)y * W;
        for (int x = roi_x; x < W; ++x) {
            const bool skipped = (!left_to_right && cs_type == 0 && l2r_cost != nullptr) &&
                                 (((x - roi_x) % D) == D - 1 || x == W - 1);
            for (int t = 0; t < D; ++t) {
                uint8_t c;
                if (skipped) {
                    c = l2r_cost[idx3(x, y, t, W, D)];
                } else {
                    const int xm = left_to_right ? x - min_d - t : x + min_d + t;
                    const uint64_t m = (xm >= 0 && xm < W) ? cm[xm] : cm[0];
                    c = (uint8_t)popcount64(cb[x] ^ m);
                }
                cost[idx3(x, y, t, W, D)] = c;
            }
        }
    }
}

/* One SGM scanline recurrence.  cusgm_ca.cu:1480-1651 (D=64 warp kernels),
 * :1654-1888 (D=128) and the generic D>=256 kernels (:156-226, :623-685,
 * :939-1005, :1891-1953) all compute, in uint32,
 *   L(p0, d) = C(p0, d)
 *   L(p, d)  = C(p, d) + min(L'(d), B-(d), B+(d), m' + P2) - m'
 *   B-(d) = d == 0   ? 127 : L'(d-1) + P1
 *   B+(d) = d == D-1 ? 127 : L'(d+1) + P1
 *   m = min_d L(p, d)
 * and store L truncated to uint8.  A8: the first pixel is stored too. */
static void scan_line(const uint8_t* cost, uint8_t* out, size_t first,
                      std::ptrdiff_t step, int n, int D, uint32_t p1, uint32_t p2,
                      std::vector<uint32_t>& Lp, std::vector<uint32_t>& Ln) {
    if (n <= 0) return;
    size_t off = first;
    uint32_t m = 0xFFFFFFFFu;
    for (int d = 0; d < D; ++d) {
        Lp[d] = cost[off + d];
        out[off + d] = (uint8_t)Lp[d];
        m = std::min(m, Lp[d]);
    }
    for (int i = 1; i < n; ++i) {
        off = (size_t)((std::ptrdiff_t)off + step);
        uint32_t mn = 0xFFFFFFFFu;
        for (int d = 0; d < D; ++d) {
            const uint32_t bm = (d == 0) ? ORC_INVALID_COST : Lp[d - 1] + p1;
            const uint32_t bp = (d == D - 1) ? ORC_INVALID_COST : Lp[d + 1] + p1;
            uint32_t v = std::min(std::min(std::min(Lp[d], bm), bp), m + p2);
            const uint32_t l = (uint32_t)cost[off + d] + (v - m);
            Ln[d] = l;
            out[off + d] = (uint8_t)l;
            mn = std::min(mn, l);
        }
        m = mn;
        std::swap(Lp, Ln);
    }
}

/* Aggregation along one direction over the region x in [roi_x, W),
 * y in [0, H) (cusgm_ca.cu:3991-4000, 4098-4108). */
void orc_aggregate(const uint8_t* cost, int W, int H, int D, int roi_x,
                   int dir, int p1, int p2, uint8_t* out) {
    const std::ptrdiff_t px = (std::ptrdiff_t)D;
    const std::ptrdiff_t row = (std::ptrdiff_t)W * D;
    const int roi_w = W - roi_x;
    if (dir == ORC_DIR_LR || dir == ORC_DIR_RL) {
#pragma omp parallel
        {
            std::vector<uint32_t> a(D), b(D);
#pragma omp for schedule(static)
            for (int y = 0; y < H; ++y) {
                if (dir == ORC_DIR_LR)
                    scan_line(cost, out, idx3(roi_x, y, 0, W, D), px, roi_w, D, p1, p2, a, b);
                else
                    scan_line(cost, out, idx3(W - 1, y, 0, W, D), -px, roi_w, D, p1, p2, a, b);
            }
        }
    } else {
#pragma omp parallel
        {
            std::vector<uint32_t> a(D), b(D);
#pragma omp for schedule(static)
            for (int x = roi_x; x < W; ++x) {
                if (dir == ORC_DIR_UD)
                    scan_line(cost, out, idx3(x, 0, 0, W, D), row, H, D, p1, p2, a, b);
                else
                    scan_line(cost, out, idx3(x, H - 1, 0, W, D), -row, H, D, p1, p2, a, b);
            }
        }
    }
}

/* Winner-takes-all with uniqueness and parabola subpixel.
 * cusgm_ca.cu:3606-3699 (D=64), :3701-3809 (D=128), :3149-3232 (generic),
 * SubPixBias :150-154.
 *   S(d) = sum of the four stored uint8 path costs
 *   border x < 5 || x > W-5 || y < 5 || y > H-5 -> INVALID
 *   c2 = min S, d* = argmin (A1: smallest d)
 *   D in {64,128}: INVALID iff d* <= 0; at d* = D-1 the reference reads c3
 *     from the shared slot that holds the argmin, i.e. c3 = (float)(D-1)
 *   other D: INVALID iff d* <= 0 || d* >= D-1
 *   uniq = min(c1 - c2, c3 - c2); INVALID iff uniq <= c2 * (1 - uniqueness)
 *   disp = (d* + min_d) + (c1 - c3 + den) / (2 den),
 *          den = max(1, c1 + c3 - 2 c2)
 * A5: the D=64/128 kernels cover rows [0, 2*(H/2)) and the generic kernel
 * all rows; columns below roi_x are never written.  Uncovered -> INVALID. */
/* WTA of a block of Wc x Hc pixels whose (0, 0) is image pixel (x0, y0) of
 * a W x H image (the whole image, or the rectangle of a region of interest):
 * the border and disparity rules use image coordinates, the covered rows
 * (2 (Hc/2) for D = 64 / 128, the reference's 2-row WTA grid over the
 * region's height) are counted from the block's first row, and block columns
 * below roi_x are INVALID.  disp is the image map (row pitch W). */
static void wta_block(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2,
                      const uint8_t* p3, int Wc, int Hc, int D, int min_d, int roi_x, int x0,
                      int y0, int W, int H, float uniqueness, float* disp) {
    const bool warp_rule = (D == 64 || D == 128);
    const int rows_covered = warp_rule ? 2 * (Hc / 2) : Hc;
    const float unique_threshold = 1.0f - uniqueness;
#pragma omp parallel
    {
        std::vector<uint32_t> S(D);
#pragma omp for schedule(static)
        for (int y = 0; y < Hc; ++y) {
            for (int x = 0; x < Wc; ++x) {
                const int xi = x + x0, yi = y + y0;
                float& out = disp[idx2(xi, yi, W)];
                out = ORC_INVALID_DISP;
                if (x < roi_x || y >= rows_covered) continue;
                if (xi > W - 5 || yi > H - 5 || xi < 5 || yi < 5) continue;
                const size_t base = idx3(x, y, 0, Wc, D);
                uint32_t c2 = 0xFFFFFFFFu;
                int dstar = 0;
                for (int d = 0; d < D; ++d) {
                    S[d] = (uint32_t)p0[base + d] + p1[base + d] + p2[base + d] + p3[base + d];
                    if (S[d] < c2) { c2 = S[d]; dstar = d; }
                }
                if (dstar <= 0) continue;
                if (!warp_rule && dstar >= D - 1) continue;
                const float c1 = (float)S[dstar - 1];
                const float c2f = (float)c2;
                const float c3 = (dstar == D - 1) ? (float)(D - 1) : (float)S[dstar + 1];
                const float a = c1 - c2f, b = c3 - c2f;
                const float uniq = (a < b) ? a : b;
                if (uniq <= c2f * unique_threshold) continue;
                const float t = c1 + c3 - 2.0f * c2f;
                const float den = (1.0f < t) ? t : 1.0f;
                const float bias = (c1 - c3 + den) / (2.0f * den);
                out = (float)(dstar + min_d) + bias;
            }
        }
    }
}

void orc_wta(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2,
             const uint8_t* p3, int W, int H, int D, int min_d, int roi_x,
             float uniqueness, float* disp) {
    wta_block(p0, p1, p2, p3, W, H, D, min_d, roi_x, 0, 0, W, H, uniqueness, disp);
}

/* Left-right consistency check.  cusgm_ca.cu:3873-3898.
 *   xr = round-half-away(xl - dL) (the +-0.5 is a double add)
 *   xr outside [0, W)      -> L INVALID
 *   |dL - dR[xr]| > thr    -> L INVALID and R[xr] INVALID
 *   otherwise              -> L = (dL + dR[xr]) / 2
 * A2: dR is read from disp_r_in (the map before the check); invalidations
 * of the right map go to disp_r_out (may be NULL).  A7: xl in [0, W). */
void orc_lrcheck(float* disp_l, const float* disp_r_in, float* disp_r_out,
                 int W, int H, float threshold) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        float* left = disp_l + (size_t)y * W;
        const float* right = disp_r_in + (size_t)y * W;
        for (int xl = 0; xl < W; ++xl) {
            const float dl = left[xl];
            if (dl == ORC_INVALID_DISP) continue;
            const float xrf = (float)xl - dl;
            const int xr = (xrf > 0) ? (int)((double)xrf + 0.5) : (int)((double)xrf - 0.5);
            if (xr < 0 || xr >= W) { left[xl] = ORC_INVALID_DISP; continue; }
            const float dr = right[xr];
            if (std::fabs(dl - dr) > threshold) {
                left[xl] = ORC_INVALID_DISP;
                if (disp_r_out) disp_r_out[(size_t)y * W + xr] = ORC_INVALID_DISP;
            } else {
                left[xl] = (dl + dr) / 2.0f;
            }
        }
    }
}

/* 3x3 median.  cusgm_df.cu:16-75 (sort network; the value it returns is the
 * exact median of the nine inputs), launch :346-355.  A3/A9: computed out of
 * place over every row; x = 0 and y = 0 copy the input; neighbours outside
 * the image count as INVALID.  The median is taken here by sorting, which is
 * independent of the network the GPU uses. */
void orc_median3x3(const float* in, float* out, int W, int H) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            if (x == 0 || y == 0) { out[idx2(x, y, W)] = in[idx2(x, y, W)]; continue; }
            float v[9];
            int k = 0;
            for (int i = -1; i <= 1; ++i)
                for (int j = -1; j <= 1; ++j) {
                    const int xx = x + j, yy = y + i;
                    v[k++] = (xx < W && yy < H) ? in[idx2(xx, yy, W)] : ORC_INVALID_DISP;
                }
            std::sort(v, v + 9);
            out[idx2(x, y, W)] = v[4];
        }
    }
}

/* ------------------------------------------------------------ pyramid --- */

/* FasterStereoCuda.cpp:114-136 (Zoom by 2): dst[i][j] = src[2i][2j]. */
void orc_zoom_half(const uint8_t* src, int w, int h, uint8_t* dst) {
    const int wd = w / 2, hd = h / 2;
    for (int i = 0; i < hd; ++i)
        for (int j = 0; j < wd; ++j) dst[(size_t)i * wd + j] = src[(size_t)(2 * i) * w + 2 * j];
}

/* FasterStereoCuda.cpp:138-166.  next is w x h; coarse is (w/2) x (h/2).
 * Even rows (but the last): b = (short)(2 d - LR/2) from the coarse pixel
 * (i/2, j/2), INVALID -> -32768; odd rows and the last row copy the row
 * above.  A18: the coarse column is clamped to the coarse width. */
void orc_init_next(const float* coarse, int w, int h, int level_range, int16_t* next) {
    const int half = level_range >> 1;
    const int wc = w >> 1;
    for (int i = 0; i < h; ++i) {
        int16_t* row = next + (size_t)i * w;
        for (int j = 0; j < w; ++j) {
            if (i % 2 == 1 || i == h - 1) {
                row[j] = next[(size_t)(i - 1) * w + j];
            } else {
                const float d = coarse[(size_t)(i >> 1) * wc + std::min(j >> 1, wc - 1)];
                row[j] = (d != ORC_INVALID_DISP) ? (int16_t)(2 * d - half) : (int16_t)ORC_INVALID_SHORT;
            }
        }
    }
}

/* cusgm_cc.cu:14-21: (uint8)expf(4*|t - LR/2| / (LR/2)) with integer
 * division, so the exponent is an integer 0..4: e^k truncated. */
static uint8_t cost_adjust(int t, int level_range) {
    static const uint8_t table[5] = {1, 2, 7, 20, 54};   // floor(e^k), k = 0..4
    const int half = level_range / 2;
    const int k = 4 * std::abs(t - half) / half;
    return table[std::min(k, 4)];
}

/* Cost with initial disparity, left view (cusgm_cc.cu:173-237 for 9x7,
 * :294-360 for 5x5): for t in [0, LR): xr = x - t - b(x);
 * 0 < xr < W -> popc(cl[x] ^ cr[xr]) + CostAdjust(t); otherwise 0xFF (9x7)
 * or 127 (5x5). */
void orc_cost_init(const uint64_t* cl, const uint64_t* cr, const int16_t* init, int W, int H,
                   int level_range, int cs_type, uint8_t* cost) {
    const int D = level_range;
    const uint8_t invalid = cs_type ? 0xFF : (uint8_t)ORC_INVALID_COST;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const int b = init[(size_t)y * W + x];
            for (int t = 0; t < D; ++t) {
                const int xr = x - t - b;
                uint8_t c = invalid;
                if (xr > 0 && xr < W)
                    c = (uint8_t)(popcount64(cl[(size_t)y * W + x] ^ cr[(size_t)y * W + xr]) +
                                  cost_adjust(t, level_range));
                cost[idx3(x, y, t, W, D)] = c;
            }
        }
}

/* Aggregation over a per-pixel window (cusgm_ca.cu:1181-1339), uint32:
 *   first pixel: L = C (A8: stored), m = min C, b' = b(p0) (A11: vertical
 *     directions take b' from row 0 whichever end they start at)
 *   b(p) INVALID: m = 127, nothing else changes, nothing is stored
 *   else: i1 = t + b - 1 - b', i2 = t + b + 1 - b'
 *     B- = i1 in [0, LR) ? L'[i1] + P1 : 127;  B+ = i2 in [0, LR) ? L'[i2] + P1 : 127
 *     P2' = max(P1, P2 / max(1, |b' - b + 1|))          (A6)
 *     L[t] = C + (min(L'[t], B-, B+, m + P2') - m)     (wraps mod 2^32)
 *     store (uint8)L, m = min L, b' = b
 * Pixels with an INVALID b keep whatever the output held (the WTA never
 * reads them). */
void orc_aggregate_init(const uint8_t* cost, const int16_t* init, int W, int H,
                        int level_range, int dir, int p1, int p2, uint8_t* out) {
    const int D = level_range;
    const bool horiz = (dir == ORC_DIR_LR || dir == ORC_DIR_RL);
    const bool fwd = (dir == ORC_DIR_UD || dir == ORC_DIR_LR);
    const int nlines = horiz ? H : W, n = horiz ? W : H;
#pragma omp parallel
    {
        std::vector<uint32_t> Lp(D), Ln(D);
#pragma omp for schedule(static)
        for (int line = 0; line < nlines; ++line) {
            auto pix = [&](int s) -> size_t {
                const int i = fwd ? s : n - 1 - s;
                return horiz ? (size_t)line * W + i : (size_t)i * W + line;
            };
            size_t p = pix(0);
            uint32_t m = 0xFFFFFFFFu;
            for (int t = 0; t < D; ++t) {
                Lp[t] = cost[p * D + t];
                out[p * D + t] = (uint8_t)Lp[t];
                m = std::min(m, Lp[t]);
            }
            int last_b = horiz ? init[p] : init[(size_t)line];
            for (int s = 1; s < n; ++s) {
                p = pix(s);
                const int b = init[p];
                if (b == ORC_INVALID_SHORT) { m = ORC_INVALID_COST; continue; }
                const int div = std::max(1, std::abs(last_b - b + 1));
                const uint32_t pen2 = (uint32_t)std::max(p1, p2 / div);
                uint32_t mn = 0xFFFFFFFFu;
                for (int t = 0; t < D; ++t) {
                    const int i1 = t + b - 1 - last_b, i2 = t + b + 1 - last_b;
                    const uint32_t bm = (i1 >= 0 && i1 < D) ? Lp[i1] + (uint32_t)p1 : ORC_INVALID_COST;
                    const uint32_t bp = (i2 >= 0 && i2 < D) ? Lp[i2] + (uint32_t)p1 : ORC_INVALID_COST;
                    const uint32_t v = std::min(std::min(Lp[t], bm), std::min(bp, m + pen2));
                    const uint32_t l = (uint32_t)cost[p * D + t] + (v - m);
                    Ln[t] = l;
                    out[p * D + t] = (uint8_t)l;
                    mn = std::min(mn, l);
                }
                m = mn;
                last_b = b;
                std::swap(Lp, Ln);
            }
        }
    }
}

/* WTA with initial disparity (cusgm_ca.cu:3425-3512): border or INVALID b ->
 * INVALID; d* = smallest argmin of S (A1); INVALID unless 2 <= d* <= LR-2;
 * uniqueness; disp = (d* + b) + SubPixBias.  Rows [4*(H/4), H) are never
 * written by the reference grid (A5: INVALID). */
void orc_wta_init(const uint8_t* p0, const uint8_t* p1, const uint8_t* p2, const uint8_t* p3,
                  const int16_t* init, int W, int H, int level_range, float uniqueness,
                  float* disp) {
    const int D = level_range;
    const int rows_covered = 4 * (H / 4);
    const float unique_threshold = 1.0f - uniqueness;
#pragma omp parallel
    {
        std::vector<uint32_t> S(D);
#pragma omp for schedule(static)
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                float& out = disp[(size_t)y * W + x];
                out = ORC_INVALID_DISP;
                if (y >= rows_covered) continue;
                if (x > W - 5 || y > H - 5 || x < 5 || y < 5) continue;
                const int b = init[(size_t)y * W + x];
                if (b == ORC_INVALID_SHORT) continue;
                const size_t base = ((size_t)y * W + x) * D;
                uint32_t c2 = 0xFFFFFFFFu;
                int ds = 0;
                for (int d = 0; d < D; ++d) {
                    S[d] = (uint32_t)p0[base + d] + p1[base + d] + p2[base + d] + p3[base + d];
                    if (S[d] < c2) { c2 = S[d]; ds = d; }
                }
                if (ds <= 1 || ds >= D - 1) continue;
                const float c1 = (float)S[ds - 1], c2f = (float)c2, c3 = (float)S[ds + 1];
                const float a = c1 - c2f, bb = c3 - c2f;
                const float uniq = (a < bb) ? a : bb;
                if (uniq <= c2f * unique_threshold) continue;
                const float t = c1 + c3 - 2.0f * c2f;
                const float den = (1.0f < t) ? t : 1.0f;
                out = (float)(ds + b) + (c1 - c3 + den) / (2.0f * den);
            }
    }
}

/* cusgm_df.cu:188-218 (Kernel_Dilation, float): an INVALID pixel strictly
 * inside the border (x > r, x < W - r, y > r, y < H - r, r = (wnd-1)/2)
 * takes the maximum of its wnd x wnd window; other pixels are copied.
 * The pyramid applies it to the coarsest level with wnd = 3 / 5 / 7 for
 * min(w, h) < 360 / < 720 / otherwise (cusgm_df.cu:300-304). */
void orc_dilation(const float* in, float* out, int W, int H, int wnd) {
    const int r = (wnd - 1) / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float v = in[(size_t)y * W + x];
            if (v == ORC_INVALID_DISP && x > r && x < W - r && y > r && y < H - r)
                for (int i = -r; i <= r; ++i)
                    for (int j = -r; j <= r; ++j) v = std::max(v, in[(size_t)(y + j) * W + x + i]);
            out[(size_t)y * W + x] = v;
        }
}

int orc_dilation_window(int w, int h) {
    const int s = std::min(w, h);
    return s < 360 ? 3 : (s < 720 ? 5 : 7);
}

int orc_match_init(const uint8_t* left, const uint8_t* right, const int16_t* init, int W, int H,
                   int level_range, const orc_options* opt, float* disp) {
    if (!left || !right || !init || !disp || W <= 0 || H <= 0) return -1;
    const size_t npix = (size_t)W * H, ncell = npix * level_range;
    std::vector<uint64_t> cl(npix), cr(npix);
    std::vector<uint8_t> cost(ncell);
    std::vector<uint8_t> path[4];
    for (auto& p : path) p.assign(ncell, 0);
    orc_census(left, W, H, opt->cs_type, 0, cl.data());
    orc_census(right, W, H, opt->cs_type, 1, cr.data());
    orc_cost_init(cl.data(), cr.data(), init, W, H, level_range, opt->cs_type, cost.data());
    for (int r = 0; r < 4; ++r)
        orc_aggregate_init(cost.data(), init, W, H, level_range, r, opt->p1, opt->p2, path[r].data());
    std::vector<float> raw(npix);
    orc_wta_init(path[0].data(), path[1].data(), path[2].data(), path[3].data(), init, W, H,
                 level_range, opt->uniqueness, raw.data());
    if (opt->do_median) orc_median3x3(raw.data(), disp, W, H);
    else std::memcpy(disp, raw.data(), npix * sizeof(float));
    return 0;
}

/* Peak removal (StereoCudaImpl.cpp:32-97 per strip, 463-498 for the strips
 * and threshold): a breadth-first flood from every unlabelled valid pixel in
 * raster order, growing across 4-neighbours of the same strip whose float
 * difference to the pixel being expanded is at most 1 (the reference's
 * abs(search - base) <= 1.0f on float32 values). */
int orc_remove_peaks(float* disp, int W, int H, int strips, float ratio) {
    if (W <= 0 || H <= 0 || strips <= 0) return 0;
    const int rows = H / strips;
    if (rows == 0) return 0;
    const int threshold = (int)((float)(W * rows) * ratio);
    int removed = 0;
    std::vector<int> label((size_t)W * rows), queue;
    for (int t = 0; t < strips; ++t) {
        float* d = disp + (size_t)t * rows * W;
        std::fill(label.begin(), label.end(), 0);
        int next = 0;
        for (int i = 0; i < W * rows; ++i) {
            if (d[i] == ORC_INVALID_DISP || label[i]) continue;
            label[i] = ++next;
            queue.assign(1, i);
            for (size_t k = 0; k < queue.size(); ++k) {
                const int c = queue[k], y = c / W, x = c % W;
                const int nb[4][2] = {{y - 1, x}, {y, x - 1}, {y, x + 1}, {y + 1, x}};
                for (const auto& n : nb) {
                    if (n[0] < 0 || n[0] >= rows || n[1] < 0 || n[1] >= W) continue;
                    const int j = n[0] * W + n[1];
                    if (label[j] || d[j] == ORC_INVALID_DISP) continue;
                    if (std::fabs(d[j] - d[c]) <= 1.0f) {
                        label[j] = next;
                        queue.push_back(j);
                    }
                }
            }
            if ((int)queue.size() < threshold) {
                for (int j : queue) d[j] = ORC_INVALID_DISP;
                removed += (int)queue.size();
            }
        }
    }
    return removed;
}

/* Coarsest interval of an option-1 pyramid (FasterStereoCuda.cpp:265-281):
 * integer division by the coarsest scale, then orc_round_interval. */
void orc_disp_interval(int min_disp, int max_disp, int W, int num_layers, int* dmin, int* range) {
    const int scale = 1 << (num_layers - 1);
    *dmin = min_disp / scale;
    orc_round_interval((double)(max_disp / scale - *dmin), W, dmin, range);
}

/* FasterStereoCuda.cpp:268-281 / :377-390: range = the smallest 2^k >= span
 * with k >= 6, halved once if above 2 W (the full width); a minimum at or
 * past W moves to W - range, not below -W. */
void orc_round_interval(double span, int W, int* dmin, int* range) {
    int k = 6;
    while (std::pow(2.0, k) < span && k < 20) ++k;
    int r = 1 << k;
    if (r > 2 * W) r /= 2;
    *range = r;
    if (*dmin >= W) {
        *dmin = W - r;
        if (*dmin < -W) *dmin = -W;
    }
}

/* Coarsest interval of a depth pyramid (HierSgmCudaImpl::Init2,
 * FasterStereoCuda.cpp:373-391): the camera scaled by s = 2^(num_layers-1)
 * (CameraParam_T::Scale, sgm_types.h:75-82: principal points and focus
 * divided by s), then in float64
 *   min = baseline * focus / max_depth - x0_right + x0_left
 *   max = baseline * focus / min_depth - x0_right + x0_left
 * and the level starts at floor(min) (the width test is on the float64
 * minimum, which is the same as on its floor since W is an integer). */
int orc_depth_interval(double min_depth, double max_depth, int W, int num_layers,
                       const double cam[6], int* dmin, int* range) {
    if (!(min_depth > 0.0) || !(max_depth > min_depth) || num_layers < 1) return -1;
    const int s = 1 << (num_layers - 1);
    const double x0l = cam[0] / s, x0r = cam[2] / s, f = cam[4] / s, b = cam[5];
    const double lo = b * f / max_depth - x0r + x0l;
    const double hi = b * f / min_depth - x0r + x0l;
    *dmin = (int)std::floor(lo);
    orc_round_interval(hi - lo, W, dmin, range);
    return 0;
}

/* Depth from disparity (cusgm_util.cu:7-33): with f1 = baseline * focus and
 * f2 = x0_right - x0_left in float64, depth = (float)(f1 / ((double)d + f2));
 * the invalid value passes through.  cam = {x0_left, y0_left, x0_right,
 * y0_right, focus, baseline}. */
void orc_depth(const float* disp, size_t n, const double cam[6], float* depth) {
    const double f1 = cam[5] * cam[4], f2 = cam[2] - cam[0];
    for (size_t i = 0; i < n; ++i)
        depth[i] = disp[i] == ORC_INVALID_DISP ? ORC_INVALID_DISP : (float)(f1 / ((double)disp[i] + f2));
}

int orc_pyramid(const uint8_t* left, const uint8_t* right, int W, int H, int min_disp,
                int max_disp, int num_layers, int level_range, const orc_options* opt,
                float* disp) {
    if (num_layers < 1 || W <= 0 || H <= 0 || max_disp <= min_disp) return -1;
    int dmin = 0, range = 0;
    orc_disp_interval(min_disp, max_disp, W, num_layers, &dmin, &range);
    return orc_pyramid_at(left, right, W, H, dmin, range, num_layers, level_range, opt, disp);
}

/* The pyramid with its coarsest level over [dmin, dmin + range). */
int orc_pyramid_at(const uint8_t* left, const uint8_t* right, int W, int H, int dmin, int range,
                   int num_layers, int level_range, const orc_options* opt, float* disp) {
    if (num_layers < 1 || W <= 0 || H <= 0 || range <= 0) return -1;
    std::vector<std::vector<uint8_t>> L(num_layers), R(num_layers);
    std::vector<int> lw(num_layers), lh(num_layers);
    for (int i = 0; i < num_layers; ++i) { lw[i] = W >> i; lh[i] = H >> i; }
    L[0].assign(left, left + (size_t)W * H);
    R[0].assign(right, right + (size_t)W * H);
    for (int i = 1; i < num_layers; ++i) {
        L[i].resize((size_t)lw[i] * lh[i]);
        R[i].resize((size_t)lw[i] * lh[i]);
        orc_zoom_half(L[i - 1].data(), lw[i - 1], lh[i - 1], L[i].data());
        orc_zoom_half(R[i - 1].data(), lw[i - 1], lh[i - 1], R[i].data());
    }
    // coarsest level: plain SGM over [dmin, dmin + range)
    const int top = num_layers - 1;
    orc_options o = *opt;
    if (top > 0) o.uniqueness = 1.0f;
    std::vector<float> cur((size_t)lw[top] * lh[top]);
    // peaks go at the coarsest level, after its filters (FasterStereoCuda.cpp:182-196)
    if (top == 0) return orc_match(left, right, W, H, dmin, range, &o, disp, nullptr);
    o.do_remove_peaks = 0;
    int rc = orc_match(L[top].data(), R[top].data(), lw[top], lh[top], dmin, range, &o, cur.data(), nullptr);
    if (rc) return rc;
    {   // the coarsest level of a multi-level pyramid is dilated (FasterStereoCuda.cpp:192-197)
        std::vector<float> dil(cur.size());
        orc_dilation(cur.data(), dil.data(), lw[top], lh[top], orc_dilation_window(lw[top], lh[top]));
        cur.swap(dil);
    }
    if (opt->do_remove_peaks)
        orc_remove_peaks(cur.data(), lw[top], lh[top], opt->peak_strips, opt->peaks_ratio);
    o.do_lr_check = 0;
    for (int n = top - 1; n >= 0; --n) {
        std::vector<int16_t> init((size_t)lw[n] * lh[n]);
        orc_init_next(cur.data(), lw[n], lh[n], level_range, init.data());
        o.uniqueness = (n == 0) ? opt->uniqueness : 1.0f;
        std::vector<float> out((size_t)lw[n] * lh[n]);
        rc = orc_match_init(L[n].data(), R[n].data(), init.data(), lw[n], lh[n], level_range, &o,
                            out.data());
        if (rc) return rc;
        cur.swap(out);
    }
    std::memcpy(disp, cur.data(), (size_t)W * H * sizeof(float));
    return 0;
}

/* Plain single-level pipeline in StereoCudaImpl::Match order
 * (StereoCudaImpl.cpp:507-546): census L, census R, cost L->R, 4 paths,
 * WTA L, [cost R->L, 4 paths, WTA R, LR check], median. */
int orc_match(const uint8_t* left, const uint8_t* right, int W, int H,
              int min_d, int D, const orc_options* opt, float* disp_l,
              float* disp_r) {
    if (!left || !right || !disp_l || !opt || W <= 0 || H <= 0 || D <= 0) return -1;
    const size_t npix = (size_t)W * H;
    const size_t ncell = npix * (size_t)D;
    const int roi_x = std::max(0, min_d);
    std::vector<uint64_t> csl(npix), csr(npix);
    std::vector<uint8_t> cost(ncell), l2r;
    std::vector<uint8_t> path[4];
    for (auto& p : path) p.assign(ncell, 0);
    std::vector<float> dl(npix), dr(npix, ORC_INVALID_DISP);

    orc_census(left, W, H, opt->cs_type, 0, csl.data());
    orc_census(right, W, H, opt->cs_type, 1, csr.data());
    orc_cost(csl.data(), csr.data(), W, H, min_d, D, 1, opt->cs_type, nullptr, cost.data());
    for (int r = 0; r < 4; ++r)
        orc_aggregate(cost.data(), W, H, D, roi_x, r, opt->p1, opt->p2, path[r].data());
    orc_wta(path[0].data(), path[1].data(), path[2].data(), path[3].data(), W, H, D, min_d,
            roi_x, opt->uniqueness, dl.data());
    if (opt->do_lr_check) {
        if (opt->cs_type == 0) l2r = cost;
        orc_cost(csr.data(), csl.data(), W, H, min_d, D, 0, opt->cs_type,
                 opt->cs_type == 0 ? l2r.data() : nullptr, cost.data());
        for (int r = 0; r < 4; ++r)
            orc_aggregate(cost.data(), W, H, D, roi_x, r, opt->p1, opt->p2, path[r].data());
        std::vector<float> dr_in(npix);
        orc_wta(path[0].data(), path[1].data(), path[2].data(), path[3].data(), W, H, D, min_d,
                roi_x, opt->uniqueness, dr_in.data());
        dr = dr_in;
        orc_lrcheck(dl.data(), dr_in.data(), dr.data(), W, H, opt->lr_threshold);
    }
    if (opt->do_median) {
        orc_median3x3(dl.data(), disp_l, W, H);
    } else {
        std::memcpy(disp_l, dl.data(), npix * sizeof(float));
    }
    // after the filters, on the left map only (StereoCudaImpl.cpp:541-542)
    if (opt->do_remove_peaks) orc_remove_peaks(disp_l, W, H, opt->peak_strips, opt->peaks_ratio);
    if (disp_r) std::memcpy(disp_r, dr.data(), npix * sizeof(float));
    return 0;
}

/* The region of interest of one view (StereoROI_T x, y, w, h; NULL = the
 * whole matched region), clipped to the image and to the columns that have a
 * cost (x >= max(0, min_d)).  Returns false when nothing is left. */
static bool roi_rect(const int* roi, int W, int H, int min_d, int* x0, int* y0, int* x1,
                     int* y1) {
    const int lo = std::max(0, min_d);
    if (!roi) {
        *x0 = lo; *y0 = 0; *x1 = W; *y1 = H;
    } else {
        *x0 = std::max(roi[0], lo);
        *y0 = std::max(roi[1], 0);
        *x1 = (int)std::min<long long>((long long)roi[0] + roi[2], W);
        *y1 = (int)std::min<long long>((long long)roi[1] + roi[3], H);
    }
    return *x0 < *x1 && *y0 < *y1;
}

/* One view restricted to a rectangle: its four directions run over the
 * rectangle only (each path's first pixel on the rectangle's edge), the WTA
 * fills the rectangle (image-coordinate border rules, rows counted from its
 * top) and every other pixel of disp is INVALID. */
static void match_rect(const uint8_t* cost, int W, int H, int D, int min_d, int x0, int y0,
                       int x1, int y1, const orc_options* opt, float* disp) {
    for (size_t i = 0; i < (size_t)W * H; ++i) disp[i] = ORC_INVALID_DISP;
    const int wc = x1 - x0, hc = y1 - y0;
    const size_t ncell = (size_t)wc * hc * D;
    std::vector<uint8_t> sub(ncell);
    for (int y = 0; y < hc; ++y)
        std::memcpy(&sub[idx3(0, y, 0, wc, D)], &cost[idx3(x0, y + y0, 0, W, D)], (size_t)wc * D);
    std::vector<uint8_t> path[4];
    for (int r = 0; r < 4; ++r) {
        path[r].assign(ncell, 0);
        orc_aggregate(sub.data(), wc, hc, D, 0, r, opt->p1, opt->p2, path[r].data());
    }
    wta_block(path[0].data(), path[1].data(), path[2].data(), path[3].data(), wc, hc, D, min_d,
              0, x0, y0, W, H, opt->uniqueness, disp);
}

/* StereoCudaImpl::Match with regions of interest (StereoCudaImpl.cpp:418-446,
 * 463-498; cusgm_ca.cu:3984-4000 left, 4180-4189 right).  The costs are
 * those of the whole image; each view's directions and WTA are restricted to
 * its rectangle (roi_l for the left view, roi_r for the right view of the LR
 * check; NULL = the whole matched region); the LR check and the median run
 * over the whole maps; peak removal splits the left rectangle's rows into
 * strips.  With both NULL this is orc_match.  A21: outside a rectangle the
 * map is INVALID (the reference leaves whatever its buffer held). */
int orc_match_roi(const uint8_t* left, const uint8_t* right, int W, int H, int min_d, int D,
                  const orc_options* opt, const int* roi_l, const int* roi_r, float* disp_l,
                  float* disp_r) {
    if (!left || !right || !disp_l || !opt || W <= 0 || H <= 0 || D <= 0) return -1;
    const size_t npix = (size_t)W * H;
    const size_t ncell = npix * (size_t)D;
    std::vector<uint64_t> csl(npix), csr(npix);
    std::vector<uint8_t> cost(ncell), l2r;
    std::vector<float> dl(npix, ORC_INVALID_DISP), dr(npix, ORC_INVALID_DISP);
    orc_census(left, W, H, opt->cs_type, 0, csl.data());
    orc_census(right, W, H, opt->cs_type, 1, csr.data());
    orc_cost(csl.data(), csr.data(), W, H, min_d, D, 1, opt->cs_type, nullptr, cost.data());
    int lx0, ly0, lx1, ly1;
    const bool left_any = roi_rect(roi_l, W, H, min_d, &lx0, &ly0, &lx1, &ly1);
    if (left_any) match_rect(cost.data(), W, H, D, min_d, lx0, ly0, lx1, ly1, opt, dl.data());
    if (opt->do_lr_check) {
        if (opt->cs_type == 0) l2r = cost;
        orc_cost(csr.data(), csl.data(), W, H, min_d, D, 0, opt->cs_type,
                 opt->cs_type == 0 ? l2r.data() : nullptr, cost.data());
        std::vector<float> dr_in(npix, ORC_INVALID_DISP);
        int rx0, ry0, rx1, ry1;
        if (roi_rect(roi_r, W, H, min_d, &rx0, &ry0, &rx1, &ry1))
            match_rect(cost.data(), W, H, D, min_d, rx0, ry0, rx1, ry1, opt, dr_in.data());
        dr = dr_in;
        orc_lrcheck(dl.data(), dr_in.data(), dr.data(), W, H, opt->lr_threshold);
    }
    if (opt->do_median) {
        orc_median3x3(dl.data(), disp_l, W, H);
    } else {
        std::memcpy(disp_l, dl.data(), npix * sizeof(float));
    }
    if (opt->do_remove_peaks && left_any)
        orc_remove_peaks(disp_l + (size_t)ly0 * W, W, ly1 - ly0, opt->peak_strips, opt->peaks_ratio);
    if (disp_r) std::memcpy(disp_r, dr.data(), npix * sizeof(float));
    return 0;
}

}  // extern "C"
