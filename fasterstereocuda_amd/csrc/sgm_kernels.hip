// SGM stereo stage kernels for CDNA4 (gfx950).
//
// Stages (one launch each, covering every job of a plan):
//   census      LDS-staged image tile -> 64-bit census per pixel
//   cost        LDS-staged census rows -> u8 Hamming cost volume [H][W][D]
//   aggregate   4-direction SGM recurrence, one wave = 4 scanlines
//   wta         4-path sum, argmin, uniqueness, parabola subpixel
//   post        LR consistency check + 3x3 median
//
// Aggregation layout: a wave64 is four 16-lane DPP rows; each row owns one
// scanline and each lane owns D/16 consecutive disparities, held as packed
// u16 pairs (v_pk_add_u16 / v_pk_sub_u16 / v_pk_min_u16).  The d-1 / d+1
// neighbours crossing a lane boundary come from row_shr:1 / row_shl:1 DPP
// moves whose "old" operand supplies the reference's INVALID_COST=127
// boundary term; the per-step minimum is a 4-step row_ror DPP reduction,
// so every lane of the row holds it without LDS or barriers.
//
// Arithmetic matches the reference exactly (ethan-li-coding/FasterStereoCuda,
// SgmStereoCuda/cusgm_cc.cu, cusgm_ca.cu, cusgm_df.cu); see each kernel.
#include <algorithm>

#include "sgm_device.h"
#include "sgm_internal.h"

#pragma clang fp contract(off)

namespace fsgm {
namespace {

using namespace dev;

// ---------------------------------------------------------------- census ---
// (acc >> 1) | (d & m) as exactly v_lshrrev + v_and_or: written in C, the
// compiler distributes the shifts of a chain of these over its ORs and
// issues about twice the instructions.
__device__ __forceinline__ uint32_t shift_in(uint32_t acc, uint32_t d, uint32_t m) {
    uint32_t r;
    asm("v_lshrrev_b32 %0, 1, %1\n\t"
        "v_and_or_b32 %0, %2, %3, %0"
        : "=&v"(r)
        : "v"(acc), "v"(d), "s"(m));
    return r;
}

// cusgm_cc.cu:30-106.  Bit = (neighbour < centre), MSB-first raster order,
// centre included; border pixels get 0 (left view) or 0xFFFFFFFF (right).
// A thread computes two horizontally adjacent pixels as the u16 halves of one
// register, for 4 vertically adjacent rows that share the window rows they
// overlap in.  The tile is staged in LDS as overlapping pairs, P[r][c] =
// (I[r][c], I[r][c + 1]), so one dword read gives a window column for both
// pixels.  Per window bit and pixel pair: d = P - centre pair (v_pk_sub_u16,
// bit 15 of each half set iff neighbour < centre), then acc = (acc >> 1) |
// (d & 0x80008000): the compare bits enter at the top of each half.  The bits
// run from the last to the first, 16 per accumulator, so each half ends up
// holding 16 consecutive bits of its pixel's word in MSB-first order (a
// half never holds more than 16, so the 32-bit shift never carries a bit
// from one half into the other).  The 64-bit word of a pixel: its last 32
// bits in the low word, the first NB - 32 in the high word.
template <int RW, int RH>
__device__ __forceinline__ void census_pairs4(const uint32_t (*t)[72], int r0, int c,
                                              uint64_t (&out)[4][2]) {
    constexpr int NJ = 2 * RW + 1, NB = NJ * (2 * RH + 1);
    constexpr int NC = (NB + 15) / 16;   // 16-bit chunks per pixel
    uint32_t acc[4][4], ctr[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        ctr[p] = t[r0 + p][c];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[p][q] = 0u;
    }
#pragma unroll
    for (int k = 3 + 2 * RH; k >= 0; --k) {       // tile rows r0 - RH .. r0 + 3 + RH, last first
        uint32_t nb[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) nb[j] = t[r0 - RH + k][c - RW + j];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const int i = k - p;                  // window row of pixel p
            if (i < 0 || i > 2 * RH) continue;
#pragma unroll
            for (int j = NJ - 1; j >= 0; --j) {
                const int q = (NB - 1 - (i * NJ + j)) / 16;   // chunk of this bit
                const uint32_t d = pk_sub(nb[j], ctr[p]);   // bit 15 of a half set iff nb < ctr
                acc[p][q] = shift_in(acc[p][q], d, 0x80008000u);
            }
        }
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        // a partial last chunk: move its bits down to the bottom of each half
        constexpr int last = NB - 16 * (NC - 1);
        if constexpr (last < 16) acc[p][NC - 1] >>= (16 - last);
        // chunks 0, 1: the low word; chunks 2, 3: the high word
        const uint32_t lo0 = __builtin_amdgcn_perm(acc[p][1], acc[p][0], 0x05040100u);
        const uint32_t lo1 = __builtin_amdgcn_perm(acc[p][1], acc[p][0], 0x07060302u);
        const uint32_t hi0 = __builtin_amdgcn_perm(acc[p][3], acc[p][2], 0x05040100u);
        const uint32_t hi1 = __builtin_amdgcn_perm(acc[p][3], acc[p][2], 0x07060302u);
        out[p][0] = ((uint64_t)hi0 << 32) | lo0;
        out[p][1] = ((uint64_t)hi1 << 32) | lo1;
    }
}

// 64 x 16 pixels per workgroup of 128 threads: thread (lx, ly) takes columns
// x0 + 2 lx, x0 + 2 lx + 1 and rows y0 + 4 ly .. y0 + 4 ly + 3.
__global__ __launch_bounds__(128) void census_kernel(const SgmJob* __restrict__ jobs,
                                                     const int* __restrict__ offs, int njobs,
                                                     IoBind io) {
    __shared__ uint8_t tile[22][72];
    __shared__ uint32_t pairs[22][72];
    const int item = blockIdx.x;
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const uint8_t* const img = (io.on && j < 2) ? io.img[j] : J.img;
    const int local = item - offs[j];
    const int tiles_x = (J.W + 63) / 64;
    const int x0 = (local % tiles_x) * 64, y0 = (local / tiles_x) * 16;
    const int tid = threadIdx.x;
    // 22 x 72 bytes over 128 threads: every load is issued before the first
    // LDS store, so the loads of a thread overlap
    constexpr int kTileN = 22 * 72, kPer = (kTileN + 127) / 128;
    uint8_t px[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = tid + 128 * u;
        const int r = idx / 72, c = idx % 72;
        const int yy = min(max(y0 - 3 + r, 0), J.H - 1);
        const int xx = min(max(x0 - 4 + c, 0), J.W - 1);
        px[u] = (idx < kTileN) ? gld(img + (int64_t)yy * J.W + xx) : (uint8_t)0;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = tid + 128 * u;
        if (idx < kTileN) tile[idx / 72][idx % 72] = px[u];
    }
    __syncthreads();
    // the overlapping pairs (column 71 has no right neighbour and is not read)
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = tid + 128 * u;
        const int r = idx / 72, c = idx % 72;
        if (idx < kTileN && c < 71) pairs[r][c] = (uint32_t)tile[r][c] | ((uint32_t)tile[r][c + 1] << 16);
    }
    __syncthreads();
    const int lx = tid & 31, ly = tid >> 5;
    const int rw = J.cs_type ? 4 : 2, rh = J.cs_type ? 3 : 2;
    const uint64_t invalid = J.is_right ? 0xFFFFFFFFull : 0ull;
    uint64_t v[4][2];
    if (J.cs_type) census_pairs4<4, 3>(pairs, ly * 4 + 3, 2 * lx + 4, v);
    else census_pairs4<2, 2>(pairs, ly * 4 + 3, 2 * lx + 4, v);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
        const int y = y0 + ly * 4 + rr;
        if (y >= J.H) continue;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int x = x0 + 2 * lx + h;
            if (x >= J.W) continue;
            const bool border = x < rw || x >= J.W - rw || y < rh || y >= J.H - rh;
            gst(J.census + (int64_t)y * J.W + x, border ? invalid : v[rr][h]);
        }
    }
}

// ------------------------------------------------------------------ cost ---
// cusgm_cc.cu:118-172 (u64) / :238-293 (u32), launch :472-523.
//   left view:  C[y][x][t] = popc(cl[x] ^ cr[x - min_d - t])
//   right view: C[y][x][t] = popc(cr[x] ^ cl[x + min_d + t])
// a partner column outside [0, W) uses column 0 of the partner census.
// 5x5 right view: the reference loop skips the last pixel of every D-pixel
// block (cusgm_cc.cu:288); its reused buffer still holds the left-view cost
// of that pixel, which is what is written here.
__global__ __launch_bounds__(256) void cost_kernel(const SgmJob* __restrict__ jobs,
                                                   const int* __restrict__ offs, int njobs) {
    __shared__ uint64_t sb[kCostSeg];
    __shared__ uint64_t sw[kCostSeg + kMaxD];
    const int item = blockIdx.x;
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const int local = item - offs[j];
    const int W = J.W, D = J.D, min_d = J.min_d, roi_x = J.roi_x;
    const int segs = (W - roi_x + kCostSeg - 1) / kCostSeg;
    const int y = local / segs;
    const int x0 = roi_x + (local - y * segs) * kCostSeg;
    const uint64_t* cb = J.census + (int64_t)y * W;
    const uint64_t* cm = J.census_other + (int64_t)y * W;
    const bool l2r = !J.is_right;
    const bool quirk_job = J.is_right && J.cs_type == 0;
    const int tid = threadIdx.x;
    sb[tid] = (x0 + tid < W) ? gld(cb + x0 + tid) : 0ull;
    const int wbase = l2r ? x0 - min_d - (D - 1) : x0 + min_d;
    const uint64_t cm0 = gld(cm);
    for (int k = tid; k < kCostSeg + D - 1; k += 256) {
        const int xm = wbase + k;
        sw[k] = (xm >= 0 && xm < W) ? gld(cm + xm) : cm0;
    }
    __syncthreads();
    // D is a power of two (64..512): groups of 16 disparities per work item
    const int lg = 31 - __builtin_clz((unsigned)D >> 4);
    const int gmask = (D >> 4) - 1;
    const int nx = min(kCostSeg, W - x0);
    uint8_t* const row_out = J.cost + ((int64_t)y * W + x0) * D;
    for (int w = tid; w < (nx << lg); w += 256) {
        const int i = w >> lg, g = w & gmask;
        const int x = x0 + i;
        const bool quirk = quirk_job && ((((x - roi_x) & (D - 1)) == D - 1) || x == W - 1);
        uint32_t out[4] = {0u, 0u, 0u, 0u};
        if (!quirk) {
            const uint64_t b = sb[i];
            const int base = l2r ? i - g * 16 + D - 1 : i + g * 16;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const uint64_t m = l2r ? sw[base - q] : sw[base + q];
                out[q >> 2] |= (uint32_t)__popcll(b ^ m) << (8 * (q & 3));
            }
        } else {
            const uint64_t b = gld(cm + x);
            const uint64_t cb0 = gld(cb);
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const int xr = x - min_d - (g * 16 + q);
                const uint64_t m = (xr >= 0 && xr < W) ? gld(cb + xr) : cb0;
                out[q >> 2] |= (uint32_t)__popcll(b ^ m) << (8 * (q & 3));
            }
        }
        *(g_u4*)(row_out + (int64_t)w * 16) = u32x4{out[0], out[1], out[2], out[3]};
    }
}

// ------------------------------------------------------------- aggregate ---
// One SGM direction over four scanlines (cusgm_ca.cu:1480-1651 for D=64,
// :1654-1888 for D=128, generic :156-226 etc. for D>=256), in exact integer
// arithmetic:
//   L(p0, d) = C(p0, d)                                   (stored, A8)
//   L(p, d)  = C(p, d) + min(L'(d), B-(d), B+(d), m' + P2) - m'
//   B-(d) = d == 0 ? 127 : L'(d-1) + P1,  B+(d) = d == D-1 ? 127 : L'(d+1) + P1
//   m = min_d L(p, d); the stored value is L mod 256 (uint8 store).
// All terms are < 2^16 when P1 + P2 <= 65000, so packed u16 is exact.
// With FUSE the direction is bottom->top and the scanline does the WTA of
// each pixel it reaches instead of storing its own volume: S = the other three
// stored directions + this one's value mod 256, argmin with the smallest d,
// c1 / c3 gathered with one row OR-reduction, then the reference's
// uniqueness and subpixel rules (see wta_pixel below).
template <int D, bool FUSE>
__device__ __forceinline__ void agg_lines(const SgmJob& J, int path, int group, int lane) {
    constexpr int DPD = D / 16;   // disparities per lane
    constexpr int NW = DPD / 4;   // u32 words of cost per lane
    constexpr int NP = DPD / 2;   // packed u16 pairs per lane
    const int sub = lane >> 4, l = lane & 15;
    const bool horiz = (path == kLR || path == kRL);
    const bool fwd = (path == kUD || path == kLR);
    const int W = J.W, H = J.H, roi_x = J.roi_x;
    const int nlines = horiz ? H : (W - roi_x);
    int line = group * 4 + sub;
    const bool active = line < nlines;
    if (!active) line = nlines - 1;
    const int n = horiz ? (W - roi_x) : H;
    int64_t x0, y0, step;
    if (horiz) {
        y0 = line; x0 = fwd ? roi_x : W - 1; step = fwd ? D : -D;
    } else {
        x0 = roi_x + line; y0 = fwd ? 0 : H - 1;
        step = fwd ? (int64_t)W * D : -(int64_t)W * D;
    }
    // Addresses = wave-uniform base (first line of the group, scalar
    // registers) + a small per-lane offset, so the per-step part of every
    // address is scalar arithmetic.
    const int line0 = group * 4;
    const int64_t offL = horiz ? ((int64_t)line0 * W + x0) * D
                               : (y0 * W + roi_x + line0) * D;
    const uint32_t voff = (uint32_t)((line - line0) * (horiz ? W * D : D) + l * DPD);
    const uint8_t* const src0 = J.cost + offL;
    uint8_t* const out_base = (path == 0) ? J.path[0] : (path == 1) ? J.path[1]
                            : (path == 2) ? J.path[2] : J.path[3];
    uint8_t* const dst0 = out_base + offL;
    const uint32_t P1v = (uint32_t)J.p1 * 0x10001u;
    const uint32_t P2v = (uint32_t)J.p2 * 0x10001u;
    const float uthr = 1.0f - J.uniqueness;
    const int min_d = J.min_d;

    uint32_t w[NW];
    load_words<NW>(src0 + voff, w);
    uint32_t L[NP];
#pragma unroll
    for (int k = 0; k < NW; ++k) { L[2 * k] = bytes01(w[k]); L[2 * k + 1] = bytes23(w[k]); }
    if (active && !FUSE) store_words<NW>(dst0 + (int64_t)0 + voff, w);
    if (FUSE && active && l == 0) gst(J.disp + y0 * W + x0, kInvalidDisp);   // row H-1: border
    uint32_t t0 = L[0];
#pragma unroll
    for (int k = 1; k < NP; ++k) t0 = pk_min(t0, L[k]);
    uint32_t m = row_min16(min(t0 & 0xFFFFu, t0 >> 16));

    // WTA of the pixel at step i from this direction's L and the other three
    // stored directions (FUSE only).  The integer part (sum, argmin with the
    // smallest d, c1 / c3 gathered with one row OR-reduction) runs every step
    // and leaves (c2 << 16 | d*, c1 << 16 | c3), equal on every lane of the
    // row; lane (i - 1) mod 16 keeps it.  Every 16 steps (and at the end of
    // the line) each lane finishes its own pixel with the float rules, so the
    // float work is done once per pixel instead of once per lane.
    uint32_t kb = 0u, kboth = 0u;
    auto wta_flush = [&](int i0, int cnt) __attribute__((always_inline)) {
        const int y = (int)y0 - (i0 + l), x = (int)x0;
        const int ds = (int)(kb & 0xFFFFu);
        const uint32_t c2 = kb >> 16;
        constexpr bool warp_rule = (D == 64 || D == 128);
        const int rows_cov = warp_rule ? 2 * (H / 2) : H;
        float outv = kInvalidDisp;
        if (y < rows_cov && !(x > W - 5 || y > H - 5 || x < 5 || y < 5) && ds > 0 &&
            (warp_rule || ds < D - 1)) {
            const float c1 = (float)(kboth >> 16);
            const float c2f = (float)c2;
            const float c3 = (ds == D - 1) ? (float)(D - 1) : (float)(kboth & 0xFFFFu);
            const float a = c1 - c2f, b = c3 - c2f;
            const float uniq = (a < b) ? a : b;
            if (!(uniq <= c2f * uthr)) {
                const float t = c1 + c3 - 2.0f * c2f;
                const float den = (1.0f < t) ? t : 1.0f;
                const float bias = (float)((double)(c1 - c3 + den) / (double)(2.0f * den));
                outv = (float)(ds + min_d) + bias;
            }
        }
        if (active && l < cnt) gst(J.disp + (int64_t)y * W + x, outv);
    };
    auto wta_step = [&](const uint32_t (&other)[3][NW], int i) __attribute__((always_inline)) {
        // sums of at most 4 x 255 stay inside each u16 half: plain u32 adds
        uint32_t S[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int wi = k >> 1;
            const uint32_t a = (k & 1) ? bytes23(other[0][wi]) : bytes01(other[0][wi]);
            const uint32_t b = (k & 1) ? bytes23(other[1][wi]) : bytes01(other[1][wi]);
            const uint32_t c = (k & 1) ? bytes23(other[2][wi]) : bytes01(other[2][wi]);
            S[k] = a + b + c + (L[k] & 0x00FF00FFu);
        }
        uint32_t best = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t d0 = (uint32_t)(l * DPD + 2 * k);
            best = min(best, (S[k] << 16) | d0);
            best = min(best, (S[k] & 0xFFFF0000u) | (d0 + 1));
        }
        best = row_min16(best);
        const int ds = (int)(best & 0xFFFFu);
        // S(ds - 1) and S(ds + 1): one owner lane each, OR-reduced over the row
        const uint32_t k1 = (uint32_t)(ds - 1 - l * DPD), k3 = k1 + 2u;
        uint32_t w1 = S[0], w3 = S[0];
#pragma unroll
        for (int k = 1; k < NP; ++k) {
            w1 = ((k1 >> 1) == (uint32_t)k) ? S[k] : w1;
            w3 = ((k3 >> 1) == (uint32_t)k) ? S[k] : w3;
        }
        // bit-field offset (k << 4) & 31 selects the u16 half
        const uint32_t v1 = (k1 < (uint32_t)DPD) ? __builtin_amdgcn_ubfe(w1, k1 << 4, 16) : 0u;
        const uint32_t v3 = (k3 < (uint32_t)DPD) ? __builtin_amdgcn_ubfe(w3, k3 << 4, 16) : 0u;
        uint32_t both = (v1 << 16) | v3;
        both |= dpp<0x128>(0u, both);
        both |= dpp<0x124>(0u, both);
        both |= dpp<0x122>(0u, both);
        both |= dpp<0x121>(0u, both);
        if (l == ((i - 1) & 15)) { kb = best; kboth = both; }
    };

    // one recurrence step: consumes the cost words of step i, updates L, m
    auto advance = [&](const uint32_t (&c8)[NW], int i) __attribute__((always_inline)) {
        const uint32_t mv = m * 0x10001u;
        const uint32_t mp2 = pk_add(mv, P2v);
        uint32_t R[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) R[k] = pk_add(L[k], P1v);
        // neighbours across lanes; lane 0 / lane 15 of a row get the 127 term
        const uint32_t prev_last = dpp<kRowShr1>(127u << 16, R[NP - 1]);
        const uint32_t next_first = dpp<kRowShl1>(127u, R[0]);
        uint32_t E[NP + 1];
        E[0] = __builtin_amdgcn_alignbit(R[0], prev_last, 16);
#pragma unroll
        for (int k = 1; k < NP; ++k) E[k] = __builtin_amdgcn_alignbit(R[k], R[k - 1], 16);
        E[NP] = __builtin_amdgcn_alignbit(next_first, R[NP - 1], 16);
        uint32_t t = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t c = (k & 1) ? bytes23(c8[k >> 1]) : bytes01(c8[k >> 1]);
            const uint32_t b = pk_min(E[k], E[k + 1]);
            const uint32_t v = pk_min(pk_min(L[k], b), mp2);
            L[k] = pk_sub(pk_add(c, v), mv);
            t = pk_min(t, L[k]);
        }
        if constexpr (!FUSE) {
            uint32_t o[NW];
#pragma unroll
            for (int k = 0; k < NW; ++k) o[k] = pack4(L[2 * k], L[2 * k + 1]);
            if (active) store_words<NW>(dst0 + (int64_t)i * step + voff, o);
        }
        m = row_min16(min(t & 0xFFFFu, t >> 16));
    };

    // Register ring of prefetched words (the cost, and with FUSE the three
    // other directions), about 64 bytes per lane in flight so a wave covers
    // many steps of memory latency.  Slot p holds step base + p; refills are
    // unconditional, clamped to the last step.
    constexpr int NSRC = FUSE ? 4 : 1;
    constexpr int PF = FUSE ? ((NW >= 4) ? 2 : 4 / NW) : ((NW >= 8) ? 2 : 16 / NW);
    const uint8_t* srcs[4] = {src0, J.path[0] + offL, J.path[1] + offL, J.path[3] + offL};
    uint32_t buf[PF][NSRC][NW];
    // The prologue issues the slots in the loop's own order (the scheduler
    // may not interleave them): the wait counts at the loop head are the
    // merge of both paths, and any other order collapses them to zero.
#pragma unroll
    for (int p = 0; p < PF; ++p) {
#pragma unroll
        for (int q = 0; q < NSRC; ++q)
            load_words<NW>(srcs[q] + (int64_t)min(1 + p, n - 1) * step + voff, buf[p][q]);
        __builtin_amdgcn_sched_barrier(0);
    }
    auto consume = [&](uint32_t (&cur)[NSRC][NW], int i) __attribute__((always_inline)) {
        advance(cur[0], i);
        if constexpr (FUSE) {
            uint32_t other[3][NW];
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 0; k < NW; ++k) other[q][k] = cur[q + 1][k];
            wta_step(other, i);
        }
    };
    // Steps i .. i + PF - 1, refilling each slot with step + PF (clamped to
    // step n - 1 near the end of the line).  Without FUSE the refill offsets
    // come from one multiply per turn and scalar adds / selects, so each load
    // adds a scalar offset to the lane's loop-invariant address instead of a
    // 64-bit multiply-add per load.  The fused kernel keeps the multiply: the
    // scalar form gives it four 64-bit lane bases and 37 more VGPRs, and its
    // larger plans run slower at the lower occupancy.
    const int64_t last_off = (int64_t)(n - 1) * step;
    auto ring_turn = [&](int i) __attribute__((always_inline)) {
        int64_t so = (int64_t)(i + PF) * step;
#pragma unroll
        for (int p = 0; p < PF; ++p) {
            uint32_t cur[NSRC][NW];
#pragma unroll
            for (int q = 0; q < NSRC; ++q)
#pragma unroll
                for (int k = 0; k < NW; ++k) cur[q][k] = buf[p][q][k];
            int64_t o;
            if constexpr (FUSE) {
                o = (int64_t)min(i + p + PF, n - 1) * step;
            } else {
                o = (i + p + PF <= n - 1) ? so : last_off;
                asm volatile("" : "+s"(o));   // keep it a scalar of its own
            }
#pragma unroll
            for (int q = 0; q < NSRC; ++q) load_words<NW>(srcs[q] + o + voff, buf[p][q]);
            so += step;
            consume(cur, i + p);
        }
    };
    int i = 1;
    if constexpr (FUSE) {
        // blocks of 16 steps (steps 16b + 1 .. 16b + 16), then their pixels
        constexpr int kUnroll = (D <= 128) ? 16 / PF : 1;
        for (; i + 16 <= n; i += 16) {
#pragma unroll kUnroll
            for (int s0 = 0; s0 < 16; s0 += PF) ring_turn(i + s0);
            wta_flush(i, 16);
        }
    } else {
        for (; i + PF <= n; i += PF) ring_turn(i);
    }
    for (; i < n; ++i) {
        uint32_t cur[NSRC][NW];
#pragma unroll
        for (int q = 0; q < NSRC; ++q) load_words<NW>(srcs[q] + (int64_t)i * step + voff, cur[q]);
        consume(cur, i);
    }
    if constexpr (FUSE) {
        // pixels of the last, partial block of 16 steps
        const int last = (n - 2) & 15;   // slot of step n - 1
        if (n > 1 && last != 15) wta_flush(n - 1 - last, last + 1);
    }
}

// MODE 0: all four directions (stage entry point); MODE 1: left->right,
// right->left, top->bottom; MODE 2: bottom->top fused with the WTA.
template <int DMAX, int MODE>
__global__ __launch_bounds__(64) void aggregate_kernel(const SgmJob* __restrict__ jobs,
                                                       const int* __restrict__ offs, int njobs) {
    const int item = blockIdx.x;
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const int local = item - offs[j];
    const int nh = (J.H + 3) / 4, nv = (J.W - J.roi_x + 3) / 4;
    int path, group;
    if (MODE == 2) { path = kDU; group = local; }
    // long horizontal scanlines are issued first
    else if (local < nh) { path = kLR; group = local; }
    else if (local < 2 * nh) { path = kRL; group = local - nh; }
    else if (local < 2 * nh + nv) { path = kUD; group = local - 2 * nh; }
    else { path = kDU; group = local - 2 * nh - nv; }
    const int lane = threadIdx.x;
    constexpr bool F = (MODE == 2);
    switch (J.D) {
        case 64: agg_lines<64, F>(J, path, group, lane); break;
        case 128: if constexpr (DMAX >= 128) agg_lines<128, F>(J, path, group, lane); break;
        case 256: if constexpr (DMAX >= 256) agg_lines<256, F>(J, path, group, lane); break;
        case 512: if constexpr (DMAX >= 512) agg_lines<512, F>(J, path, group, lane); break;
        default: break;
    }
}

// Persistent form of MODE 1 / MODE 2: a fixed grid of resident waves pulls
// packed items (job << 16 | direction << 14 | group) from a device counter
// until the list is exhausted; every wave leaves through that one test.  The
// list order is the scanline schedule (e.g. longest scanlines first).
template <int DMAX, int MODE>
__global__ __launch_bounds__(64) void aggregate_persistent_kernel(
    const SgmJob* __restrict__ jobs, const uint32_t* __restrict__ items, int n_items,
    uint32_t* __restrict__ counter) {
    const int lane = threadIdx.x;
    constexpr bool F = (MODE == 2);
    for (;;) {
        uint32_t idx = 0;
        if (lane == 0) idx = atomicAdd(counter, 1u);
        idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(idx, 0));
        if (idx >= (uint32_t)n_items) break;
        const uint32_t it = items[idx];
        const SgmJob& J = jobs[it >> 16];
        const int path = (int)((it >> 14) & 3u), group = (int)(it & 0x3FFFu);
        switch (J.D) {
            case 64: agg_lines<64, F>(J, path, group, lane); break;
            case 128: if constexpr (DMAX >= 128) agg_lines<128, F>(J, path, group, lane); break;
            case 256: if constexpr (DMAX >= 256) agg_lines<256, F>(J, path, group, lane); break;
            case 512: if constexpr (DMAX >= 512) agg_lines<512, F>(J, path, group, lane); break;
            default: break;
        }
    }
}

__global__ void reset_counters_kernel(uint32_t* c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// ------------------------------------------------------------------- wta ---
// cusgm_ca.cu:3606-3699 (D=64), :3701-3809 (D=128), :3149-3232 (generic),
// SubPixBias :150-154.  16 lanes per pixel, D/16 disparities per lane.
// argmin ties -> smallest d (key = S << 16 | d).
template <int D>
__device__ __forceinline__ void wta_pixel(const SgmJob& J, int x, int y, int l,
                                          uint16_t* __restrict__ s) {
    constexpr int DPD = D / 16, NW = DPD / 4, NP = DPD / 2;
    const bool inimg = x < J.W;
    const int xc = inimg ? x : J.W - 1;
    const int64_t off = ((int64_t)y * J.W + xc) * D + l * DPD;
    uint32_t S[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) S[k] = 0u;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        uint32_t w[NW];
        load_words<NW>(J.path[r] + off, w);
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            S[2 * k] = pk_add(S[2 * k], bytes01(w[k]));
            S[2 * k + 1] = pk_add(S[2 * k + 1], bytes23(w[k]));
        }
    }
    uint32_t best = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const uint32_t d0 = (uint32_t)(l * DPD + 2 * k);
        best = min(best, (S[k] << 16) | d0);
        best = min(best, (S[k] & 0xFFFF0000u) | (d0 + 1));
    }
    best = row_min16(best);
    uint32_t* s32 = reinterpret_cast<uint32_t*>(s);
#pragma unroll
    for (int k = 0; k < NP; ++k) s32[l * NP + k] = S[k];
    __syncthreads();
    if (l == 0 && inimg) {
        // rows are counted in the job; the border is the image's (the frame)
        const int W = J.fW ? J.fW : J.W, H = J.fW ? J.fH : J.H;
        const int xi = x + J.fx0, yi = y + J.fy0;
        constexpr bool warp_rule = (D == 64 || D == 128);
        const int rows_cov = warp_rule ? 2 * (J.H / 2) : J.H;
        float outv = kInvalidDisp;
        if (x >= J.roi_x && y < rows_cov && !(xi > W - 5 || yi > H - 5 || xi < 5 || yi < 5)) {
            const uint32_t c2 = best >> 16;
            const int ds = (int)(best & 0xFFFFu);
            if (ds > 0 && (warp_rule || ds < D - 1)) {
                const float c1 = (float)s[ds - 1];
                const float c2f = (float)c2;
                const float c3 = (ds == D - 1) ? (float)(D - 1) : (float)s[ds + 1];
                const float a = c1 - c2f, b = c3 - c2f;
                const float uniq = (a < b) ? a : b;
                const float ut = 1.0f - J.uniqueness;
                if (!(uniq <= c2f * ut)) {
                    const float t = c1 + c3 - 2.0f * c2f;
                    const float den = (1.0f < t) ? t : 1.0f;
                    // correctly rounded single-precision quotient
                    const float bias = (float)((double)(c1 - c3 + den) / (double)(2.0f * den));
                    outv = (float)(ds + J.min_d) + bias;
                }
            }
        }
        gst(J.disp + (int64_t)yi * W + xi, outv);
    }
}

template <int DMAX>
__global__ __launch_bounds__(256) void wta_kernel(const SgmJob* __restrict__ jobs,
                                                  const int* __restrict__ offs, int njobs) {
    __shared__ __align__(16) uint16_t sS[16][DMAX];
    const int item = blockIdx.x;
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const int local = item - offs[j];
    const int segs = (J.W + 15) / 16;
    const int y = local / segs;
    const int pix = threadIdx.x >> 4, l = threadIdx.x & 15;
    const int x = (local - y * segs) * 16 + pix;
    switch (J.D) {
        case 64: wta_pixel<64>(J, x, y, l, sS[pix]); break;
        case 128: if constexpr (DMAX >= 128) wta_pixel<128>(J, x, y, l, sS[pix]); break;
        case 256: if constexpr (DMAX >= 256) wta_pixel<256>(J, x, y, l, sS[pix]); break;
        case 512: if constexpr (DMAX >= 512) wta_pixel<512>(J, x, y, l, sS[pix]); break;
        default: break;
    }
}

// ------------------------------------------------------------------ post ---
// LR check (cusgm_ca.cu:3873-3898) evaluated per pixel against the right
// map as it was before the check, then the 3x3 median (cusgm_df.cu:16-75)
// out of place; x = 0 / y = 0 pass through, outside neighbours are INVALID.
// columns below roi_x are never matched: INVALID whatever the raw map holds
__device__ __forceinline__ float raw_at(const float* m, const SgmPost& P, int x, int y) {
    return (x < P.roi_x) ? kInvalidDisp : gld(m + (int64_t)y * P.W + x);
}

// the LR rule for the left value dl of (x, y)
__device__ __forceinline__ float lr_finish(const SgmPost& P, int x, int y, float dl) {
    if (!P.do_lr || dl == kInvalidDisp) return dl;
    const float xrf = (float)x - dl;
    const int xr = (xrf > 0) ? (int)((double)xrf + 0.5) : (int)((double)xrf - 0.5);
    if (xr < 0 || xr >= P.W) return kInvalidDisp;
    const float dr = raw_at(P.raw_r, P, xr, y);
    if (fabsf(dl - dr) > P.lr_thr) return kInvalidDisp;
    return (dl + dr) / 2.0f;
}

__device__ __forceinline__ float median9(float a0, float a1, float a2, float b0, float b1,
                                         float b2, float c0, float c1, float c2) {
    // sort each column (a_i, b_i, c_i); the median of the nine values is
    // med3(max of the column minima, med of the column medians, min of the maxima)
    const float lo0 = fminf(fminf(a0, b0), c0), hi0 = fmaxf(fmaxf(a0, b0), c0);
    const float lo1 = fminf(fminf(a1, b1), c1), hi1 = fmaxf(fmaxf(a1, b1), c1);
    const float lo2 = fminf(fminf(a2, b2), c2), hi2 = fmaxf(fmaxf(a2, b2), c2);
    const float md0 = __builtin_amdgcn_fmed3f(a0, b0, c0);
    const float md1 = __builtin_amdgcn_fmed3f(a1, b1, c1);
    const float md2 = __builtin_amdgcn_fmed3f(a2, b2, c2);
    const float p = fmaxf(fmaxf(lo0, lo1), lo2);
    const float q = __builtin_amdgcn_fmed3f(md0, md1, md2);
    const float r = fminf(fminf(hi0, hi1), hi2);
    return __builtin_amdgcn_fmed3f(p, q, r);
}

// 64 x 16 pixels per workgroup of 256 threads.  The tile with its one-pixel
// halo (66 x 18 LR-checked values) is built in LDS first: a thread issues
// its 5 left-map loads, then the 5 dependent right-map loads, so a workgroup
// waits for two load latencies in all.  Each thread then takes a column of 4
// pixels, whose 3 x 3 windows share the LDS rows they overlap in.
constexpr int kPostTW = 66, kPostTH = 18;
__global__ __launch_bounds__(256) void post_kernel(const SgmPost* __restrict__ posts,
                                                   const int* __restrict__ offs, int nposts,
                                                   IoBind io) {
    __shared__ float t[kPostTH][kPostTW];
    const int item = blockIdx.x;
    const int j = find_job(offs, nposts, item);
    const SgmPost& P = posts[j];
    const bool bound = io.on && j == 0;
    float* const out_l = bound ? io.out_l : P.out_l;
    float* const out_r = bound ? io.out_r : P.out_r;
    const int local = item - offs[j];
    const int tiles_x = (P.W + 63) / 64;
    const int x0 = (local % tiles_x) * 64, y0 = (local / tiles_x) * 16;
    const int tid = threadIdx.x;
    constexpr int kN = kPostTW * kPostTH, kPer = (kN + 255) / 256;
    float dl[kPer];
    int ex[kPer], ey[kPer];
    bool in[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = tid + 256 * u;
        ex[u] = x0 - 1 + idx % kPostTW;
        ey[u] = y0 - 1 + idx / kPostTW;
        in[u] = idx < kN && ex[u] >= 0 && ey[u] >= 0 && ex[u] < P.W && ey[u] < P.H;
        dl[u] = in[u] ? raw_at(P.raw_l, P, ex[u], ey[u]) : kInvalidDisp;
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int idx = tid + 256 * u;
        if (idx < kN)
            t[idx / kPostTW][idx % kPostTW] = in[u] ? lr_finish(P, ex[u], ey[u], dl[u]) : kInvalidDisp;
    }
    __syncthreads();
    const int tx = tid & 63, ty = (tid >> 6) * 4;
    const int x = x0 + tx;
    if (x >= P.W) return;
    float c[6][3];   // tile rows ty .. ty + 5, columns tx .. tx + 2
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) c[r][k] = t[ty + r][tx + k];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int y = y0 + ty + r;
        if (y >= P.H) break;
        // the right output starts as the right map (LR invalidations follow
        // in a second launch) or, without the LR check, as all INVALID
        if (out_r)
            gst(out_r + (int64_t)y * P.W + x, P.do_lr ? raw_at(P.raw_r, P, x, y) : kInvalidDisp);
        float v = c[r + 1][1];
        if (P.do_median && x > 0 && y > 0)
            v = median9(c[r][0], c[r][1], c[r][2], c[r + 1][0], c[r + 1][1], c[r + 1][2],
                        c[r + 2][0], c[r + 2][1], c[r + 2][2]);
        gst(out_l + (int64_t)y * P.W + x, v);
    }
}

// Right-map invalidation of the LR check (cusgm_ca.cu:3894-3895): every left
// pixel that fails marks its partner INVALID in out_r, which holds a copy of
// the right map.  All writers store the same value, so the result is fixed.
// Same 64 x 16 tiles as post_kernel, a column of 4 pixels per thread.
__global__ __launch_bounds__(256) void right_invalidate_kernel(const SgmPost* __restrict__ posts,
                                                               const int* __restrict__ offs,
                                                               int nposts, IoBind io) {
    const int item = blockIdx.x;
    const int j = find_job(offs, nposts, item);
    const SgmPost& P = posts[j];
    float* const out_r = (io.on && j == 0) ? io.out_r : P.out_r;
    if (!out_r || !P.do_lr) return;
    const int local = item - offs[j];
    const int tiles_x = (P.W + 63) / 64;
    const int x = (local % tiles_x) * 64 + (threadIdx.x & 63);
    const int yb = (local / tiles_x) * 16 + (threadIdx.x >> 6) * 4;
    if (x >= P.W) return;
    float dls[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) dls[r] = (yb + r < P.H) ? raw_at(P.raw_l, P, x, yb + r) : kInvalidDisp;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int y = yb + r;
        const float dl = dls[r];
        if (y >= P.H || dl == kInvalidDisp) continue;
        const float xrf = (float)x - dl;
        const int xr = (xrf > 0) ? (int)((double)xrf + 0.5) : (int)((double)xrf - 0.5);
        if (xr < 0 || xr >= P.W) continue;
        const float dr = raw_at(P.raw_r, P, xr, y);
        if (fabsf(dl - dr) > P.lr_thr) gst(out_r + (int64_t)y * P.W + xr, kInvalidDisp);
    }
}

// Depth from disparity (cusgm_util.cu:7-33): f1 = baseline * focus and
// f2 = x0_right - x0_left, both float64; the division is IEEE float64 and the
// result is rounded once to float32.  Invalid pixels stay invalid.
__global__ __launch_bounds__(256) void depth_kernel(const float* __restrict__ disp,
                                                    float* __restrict__ depth, size_t n,
                                                    double f1, double f2) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float d = disp[i];
        depth[i] = d == kInvalidDisp ? kInvalidDisp : (float)(f1 / ((double)d + f2));
    }
}

// A rectangle of a cost volume, packed: 16 bytes per lane.
__global__ __launch_bounds__(256) void crop_volume_kernel(const uint8_t* __restrict__ src,
                                                          uint8_t* __restrict__ dst, int W, int D,
                                                          int x0, int y0, int wc, int hc) {
    const int64_t row = (int64_t)wc * D / 16;          // 16-byte chunks per packed row
    const int64_t n = row * hc;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t y = i / row, r = i - y * row;
        const uint4 v = *reinterpret_cast<const uint4*>(src + ((y + y0) * W + x0) * D + r * 16);
        *reinterpret_cast<uint4*>(dst + i * 16) = v;
    }
}

// Columns [0, ncols) of every row set to INVALID: the part of a raw map left
// of the matched region, which the fused on-the-fly WTA does not write.
__global__ __launch_bounds__(256) void invalid_columns_kernel(float* __restrict__ disp, int W,
                                                              int H, int ncols) {
    const int64_t n = (int64_t)ncols * H;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int64_t y = i / ncols, x = i - y * ncols;
        disp[y * W + x] = kInvalidDisp;
    }
}

}  // namespace

hipError_t launch_crop_volume(const uint8_t* src, uint8_t* dst, int W, int D, int x0, int y0,
                              int wc, int hc, hipStream_t s) {
    if (wc <= 0 || hc <= 0 || D % 16 != 0) return wc <= 0 || hc <= 0 ? hipSuccess : hipErrorInvalidValue;
    const int64_t n = (int64_t)wc * hc * D / 16;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(crop_volume_kernel, dim3(blocks), dim3(256), 0, s, src, dst, W, D, x0, y0,
                       wc, hc);
    return hipGetLastError();
}

hipError_t launch_invalid_columns(float* disp, int W, int H, int ncols, hipStream_t s) {
    ncols = std::min(ncols, W);
    if (ncols <= 0 || H <= 0) return hipSuccess;
    const int64_t n = (int64_t)ncols * H;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(invalid_columns_kernel, dim3(blocks), dim3(256), 0, s, disp, W, H, ncols);
    return hipGetLastError();
}

hipError_t launch_depth(const float* disp, float* depth, size_t n, double f1, double f2,
                        hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(depth_kernel, dim3(blocks), dim3(256), 0, s, disp, depth, n, f1, f2);
    return hipGetLastError();
}

hipError_t launch_census(const StagePlan& p, hipStream_t s) {
    if (p.n_census <= 0) return hipSuccess;
    hipLaunchKernelGGL(census_kernel, dim3(p.n_census), dim3(128), 0, s, p.jobs, p.offs_census,
                       p.njobs, p.io);
    return hipGetLastError();
}

hipError_t launch_cost(const StagePlan& p, hipStream_t s) {
    if (p.n_cost <= 0) return hipSuccess;
    hipLaunchKernelGGL(cost_kernel, dim3(p.n_cost), dim3(256), 0, s, p.jobs, p.offs_cost, p.njobs);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_agg_mode(const StagePlan& p, const int* offs, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(64);
    if (p.dmax <= 64)
        hipLaunchKernelGGL((aggregate_kernel<64, MODE>), g, b, 0, s, p.jobs, offs, p.njobs);
    else if (p.dmax <= 128)
        hipLaunchKernelGGL((aggregate_kernel<128, MODE>), g, b, 0, s, p.jobs, offs, p.njobs);
    else if (p.dmax <= 256)
        hipLaunchKernelGGL((aggregate_kernel<256, MODE>), g, b, 0, s, p.jobs, offs, p.njobs);
    else
        hipLaunchKernelGGL((aggregate_kernel<512, MODE>), g, b, 0, s, p.jobs, offs, p.njobs);
    return hipGetLastError();
}

hipError_t launch_aggregate(const StagePlan& p, hipStream_t s) {
    return launch_agg_mode<0>(p, p.offs_agg, p.n_agg, s);
}

hipError_t launch_aggregate3(const StagePlan& p, hipStream_t s) {
    return launch_agg_mode<1>(p, p.offs_agg3, p.n_agg3, s);
}

hipError_t launch_aggregate_wta(const StagePlan& p, hipStream_t s) {
    return launch_agg_mode<2>(p, p.offs_dw, p.n_dw, s);
}

template <int MODE>
static hipError_t launch_persistent_mode(const StagePlan& p, const uint32_t* items, int n,
                                         uint32_t* counter, int grid, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 g(grid), b(64);
    if (p.dmax <= 64)
        hipLaunchKernelGGL((aggregate_persistent_kernel<64, MODE>), g, b, 0, s, p.jobs, items, n, counter);
    else if (p.dmax <= 128)
        hipLaunchKernelGGL((aggregate_persistent_kernel<128, MODE>), g, b, 0, s, p.jobs, items, n, counter);
    else if (p.dmax <= 256)
        hipLaunchKernelGGL((aggregate_persistent_kernel<256, MODE>), g, b, 0, s, p.jobs, items, n, counter);
    else
        hipLaunchKernelGGL((aggregate_persistent_kernel<512, MODE>), g, b, 0, s, p.jobs, items, n, counter);
    return hipGetLastError();
}

hipError_t launch_reset_counters(uint32_t* counters, int n, hipStream_t s) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters, n);
    return hipGetLastError();
}

hipError_t launch_aggregate_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s) {
    hipError_t e = launch_reset_counters(sc.counters, 2, s);
    if (e != hipSuccess) return e;
    if ((e = launch_persistent_mode<1>(p, sc.items3, sc.n3, sc.counters, sc.grid, s)) != hipSuccess)
        return e;
    return hipSuccess;
}

hipError_t launch_aggregate_wta_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s) {
    return launch_persistent_mode<2>(p, sc.items_dw, sc.ndw, sc.counters + 1, sc.grid, s);
}

template <int MODE>
static hipError_t occ_mode(int dmax, int* n) {
    if (dmax <= 64)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, aggregate_persistent_kernel<64, MODE>, 64, 0);
    if (dmax <= 128)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, aggregate_persistent_kernel<128, MODE>, 64, 0);
    if (dmax <= 256)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, aggregate_persistent_kernel<256, MODE>, 64, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(n, aggregate_persistent_kernel<512, MODE>, 64, 0);
}

hipError_t persistent_occupancy(const StagePlan& p, int* per_cu) {
    int a = 0, b = 0;
    hipError_t e = occ_mode<1>(p.dmax, &a);
    if (e != hipSuccess) return e;
    if ((e = occ_mode<2>(p.dmax, &b)) != hipSuccess) return e;
    *per_cu = std::max(1, std::min(a, b));
    return hipSuccess;
}

hipError_t launch_wta(const StagePlan& p, hipStream_t s) {
    if (p.n_wta <= 0) return hipSuccess;
    const dim3 g(p.n_wta), b(256);
    if (p.dmax <= 64)
        hipLaunchKernelGGL(wta_kernel<64>, g, b, 0, s, p.jobs, p.offs_wta, p.njobs);
    else if (p.dmax <= 128)
        hipLaunchKernelGGL(wta_kernel<128>, g, b, 0, s, p.jobs, p.offs_wta, p.njobs);
    else if (p.dmax <= 256)
        hipLaunchKernelGGL(wta_kernel<256>, g, b, 0, s, p.jobs, p.offs_wta, p.njobs);
    else
        hipLaunchKernelGGL(wta_kernel<512>, g, b, 0, s, p.jobs, p.offs_wta, p.njobs);
    return hipGetLastError();
}

hipError_t launch_post(const StagePlan& p, hipStream_t s) {
    if (p.n_post <= 0) return hipSuccess;
    hipLaunchKernelGGL(post_kernel, dim3(p.n_post), dim3(256), 0, s, p.posts, p.offs_post,
                       p.nposts, p.io);
    return hipGetLastError();
}

hipError_t launch_right_invalidate(const StagePlan& p, hipStream_t s) {
    if (p.n_post <= 0) return hipSuccess;
    hipLaunchKernelGGL(right_invalidate_kernel, dim3(p.n_post), dim3(256), 0, s, p.posts,
                       p.offs_post, p.nposts, p.io);
    return hipGetLastError();
}

}  // namespace fsgm
