// Coarse-to-fine pyramid kernels for CDNA4 (gfx950): 2x decimation,
// initial-disparity propagation, and the per-pixel-window cost, aggregation
// and WTA of the finer levels (ethan-li-coding/FasterStereoCuda:
// FasterStereoCuda.cpp:114-166, SgmStereoCuda/cusgm_cc.cu:14-21, 173-237,
// 294-360, cusgm_ca.cu:1181-1339, 3425-3512).
//
// A finer level searches LR (16 or 32) disparities t around a per-pixel base
// b(p): d = b(p) + t.  Its aggregation keeps one disparity per lane (a wave
// holds 64/LR scanlines) because the previous pixel's window is shifted by
// b - b' and the neighbour terms index L'[t + b - b' -+ 1] -- an arbitrary
// lane shift, fetched with ds_bpermute.  Arithmetic is uint32 with the
// reference's wrap-around after INVALID base pixels.
#include <algorithm>
#include <atomic>
#include <mutex>

#include "pyr_internal.h"

#pragma clang fp contract(off)

namespace fsgm {
namespace {

constexpr int kInvalidShort = -32768;

__device__ __forceinline__ uint32_t bperm(int src_lane, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src_lane << 2, (int)v);
}

// lanes whose source is outside the row read 0 (bound_ctrl), so the move
// needs no zeroed destination first
template <int CTRL>
__device__ __forceinline__ uint32_t dppu(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// minimum over each aligned group of LR lanes (LR = 16 or 32)
template <int LR>
__device__ __forceinline__ uint32_t group_min(uint32_t m, int lane) {
    m = min(m, dppu<0x128>(m));
    m = min(m, dppu<0x124>(m));
    m = min(m, dppu<0x122>(m));
    m = min(m, dppu<0x121>(m));
    if constexpr (LR == 32) m = min(m, bperm(lane ^ 16, m));
    return m;
}

__global__ __launch_bounds__(256) void zoom_kernel(PyrZoom z) {
    const int wd = z.w / 2, hd = z.h / 2;
    const int64_t n = (int64_t)wd * hd;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= 2 * n) return;
    const int v = i >= n;
    const int64_t k = v ? i - n : i;
    const int y = (int)(k / wd), x = (int)(k - (int64_t)y * wd);
    const uint8_t* src = v ? z.src_r : z.src_l;
    uint8_t* dst = v ? z.dst_r : z.dst_l;
    dst[k] = src[(int64_t)(2 * y) * z.w + 2 * x];
}

// FasterStereoCuda.cpp:138-166.  Row i takes its value from row
// src_row(i): odd rows copy the row above, and the last row (when even)
// copies the row above it, which copies the one above that.
__global__ __launch_bounds__(256) void init_next_kernel(PyrInit a) {
    const int64_t n = (int64_t)a.w * a.h;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int y = (int)(i / a.w), x = (int)(i - (int64_t)y * a.w);
    int r = y;
    if (r % 2 == 1) r -= 1;            // odd row: the even row above
    else if (r == a.h - 1) r -= 2;     // even last row: row above, itself a copy of y-2
    int16_t v = kInvalidShort;
    if (r >= 0) {
        const int wc = a.w >> 1;
        const float d = a.coarse[(int64_t)(r >> 1) * wc + min(x >> 1, wc - 1)];
        if (d != kInvalidDisp) v = (int16_t)(2 * d - (a.level_range >> 1));
    }
    a.next[i] = v;
}

__device__ __forceinline__ uint32_t cost_adjust(int t, int lr) {
    // (uint8)expf(k) for the integer exponent k = 4|t - LR/2| / (LR/2) in 0..4
    const int half = lr / 2;
    const int k = min(4 * abs(t - half) / half, 4);
    return k == 0 ? 1u : k == 1 ? 2u : k == 2 ? 7u : k == 3 ? 20u : 54u;
}

// cost with initial disparity, left view: xr = x - t - b;
// 0 < xr < W -> popc + CostAdjust(t); else 0xFF (9x7) or 127 (5x5)
template <int LR>
__global__ __launch_bounds__(256) void cost_init_kernel(PyrLevel a) {
    const int64_t n = (int64_t)a.w * a.h;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int y = (int)(i / a.w), x = (int)(i - (int64_t)y * a.w);
    const int b = a.init[i];
    const uint64_t cl = a.census_l[i];
    const uint64_t* cr = a.census_r + (int64_t)y * a.w;
    const uint32_t invalid = a.cs_type ? 0xFFu : 127u;
    uint32_t words[LR / 4];
#pragma unroll
    for (int k = 0; k < LR / 4; ++k) words[k] = 0u;
    if (x - b - (LR - 1) > 0 && x - b < a.w) {
        // every partner column of the window is inside the image (all but
        // the pixels near the borders and those with an invalid base)
        const uint64_t* crb = cr + (x - b);
#pragma unroll
        for (int t = 0; t < LR; ++t) {
            const uint32_t c = (uint32_t)__popcll(cl ^ crb[-t]) + cost_adjust(t, LR);
            words[t >> 2] |= (c & 0xFFu) << (8 * (t & 3));
        }
    } else {
#pragma unroll
        for (int t = 0; t < LR; ++t) {
            const int xr = x - t - b;
            uint32_t c = invalid;
            if (xr > 0 && xr < a.w) c = (uint32_t)__popcll(cl ^ cr[xr]) + cost_adjust(t, LR);
            words[t >> 2] |= (c & 0xFFu) << (8 * (t & 3));
        }
    }
    uint4* dst = reinterpret_cast<uint4*>(a.cost + i * LR);
#pragma unroll
    for (int q = 0; q < LR / 16; ++q)
        dst[q] = make_uint4(words[4 * q], words[4 * q + 1], words[4 * q + 2], words[4 * q + 3]);
}

// Aggregation over the per-pixel window (cusgm_ca.cu:1181-1339), one
// disparity per lane, SPW = 64 / LR scanlines per wave.  Step rules, with b
// and b' the window bases of this and the previous pixel: the neighbour
// terms read the previous L at window index t + b - 1 - b' and t + b + 1 - b'
// plus P1 (127 outside [0, LR)), the same-d term is L(t) unshifted, P2 is
// max(P1, P2 / max(1, |b' - b + 1|)), and a pixel with an invalid base keeps
// the previous L and b' and sets m' = 127.  The first pixel of a vertical
// scanline takes b' from row 0.  A scanline is walked in blocks of K = LR steps.  For the loads
// and stores lane j stands for pixel (line j / K, step j % K) of a block: it
// fetches that pixel's whole window (LR cost bytes) and base disparity one
// block ahead, and the block passes through a per-wave LDS area, from which
// every step reads its byte and base.  Each step's results are gathered
// there and leave as one LR-byte store per lane and block.  (Loading one
// byte per lane and step, and waiting for it, made the scanline walk
// latency-bound.)  Pixels whose base is invalid are never read by the WTA;
// they are written as 0.
template <int LR>
struct AggInitLds {
    uint8_t cost[64 * LR];   // [line][step][t]
    uint8_t out[64 * LR];
    int16_t init[64];        // [line][step]
    uint32_t pen2[256];      // max(P1, P2 / k), k = max(1, index)
};

template <int LR, bool PEN2_TABLE>
__device__ __forceinline__ void agg_init_lines(const PyrLevel& a, int dir, int group, int lane,
                                               AggInitLds<LR>& S) {
    constexpr int SPW = 64 / LR, K = LR, NV = LR / 16;
    const int sub = lane / LR, t = lane % LR;        // recurrence: line, disparity
    const int lsub = lane / K, lstep = lane % K;     // loads / stores: line, step
    const bool horiz = (dir == 1 || dir == 3);
    const bool fwd = (dir == 0 || dir == 1);
    const int W = a.w, H = a.h;
    const int nlines = horiz ? H : W;
    const int n = horiz ? W : H;
    // pixel of (line, step); lines past the last and steps past the end are
    // clamped (computed, never stored)
    auto pix_of = [&](int line, int step) -> int64_t {
        line = min(line, nlines - 1);
        step = min(step, n - 1);
        const int pos = fwd ? step : n - 1 - step;
        return horiz ? (int64_t)line * W + pos : (int64_t)pos * W + line;
    };
    const int line = min(group * SPW + sub, nlines - 1);
    const int lline = group * SPW + lsub;
    uint8_t* out = (dir == 0) ? a.path[0] : (dir == 1) ? a.path[1] : (dir == 2) ? a.path[2] : a.path[3];
    const uint32_t p1 = (uint32_t)a.p1;
    const int gbase = sub * LR;

    static_assert(NV == 1 || NV == 2, "windows of 16 or 32 bytes");
    uint4 nc0 = make_uint4(0u, 0u, 0u, 0u), nc1 = nc0;   // named, so they stay in registers
    int nb = 0;
    auto fetch = [&](int blk) __attribute__((always_inline)) {
        const int64_t p = pix_of(lline, blk * K + lstep);
        const uint4* src = reinterpret_cast<const uint4*>(a.cost + p * LR);
        nc0 = src[0];
        if constexpr (NV == 2) nc1 = src[1];
        nb = a.init[p];
    };
    auto stage = [&]() __attribute__((always_inline)) {
        uint4* dst = reinterpret_cast<uint4*>(S.cost + lane * LR);
        dst[0] = nc0;
        if constexpr (NV == 2) dst[1] = nc1;
        S.init[lane] = (int16_t)nb;
    };

    // P2 / max(1, |b' - b + 1|) by table (PEN2_TABLE, for P2 <= 254): every
    // divisor from 255 on gives P2 / k = 0, the same as entry 255, so the
    // table is exact.  A larger P2 divides (an integer division costs ~15
    // VALU per step).
    if constexpr (PEN2_TABLE) {
#pragma unroll
        for (int k = lane; k < 256; k += 64) S.pen2[k] = (uint32_t)max(a.p1, a.p2 / max(1, k));
    }

    uint32_t L = 0, m = 0;
    int last_b = 0;
    const int nblk = (n + K - 1) / K;
    fetch(0);
    stage();
    if (nblk > 1) fetch(1);
    for (int blk = 0; blk < nblk; ++blk) {
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const uint32_t c = S.cost[(sub * K + s) * LR + t];
            const int b = S.init[sub * K + s];
            uint32_t o;
            if (s == 0 && blk == 0) {
                L = c;
                m = group_min<LR>(L, lane);
                // A11: vertical directions seed b' from row 0 whichever end they start at
                last_b = horiz ? b : a.init[line];
                o = L;
            } else {
                const bool valid = b != kInvalidShort;
                const int i1 = t + b - 1 - last_b, i2 = t + b + 1 - last_b;
                const uint32_t v1 = bperm(gbase + min(max(i1, 0), LR - 1), L);
                const uint32_t v2 = bperm(gbase + min(max(i2, 0), LR - 1), L);
                const uint32_t bm = (i1 >= 0 && i1 < LR) ? v1 + p1 : 127u;
                const uint32_t bp = (i2 >= 0 && i2 < LR) ? v2 + p1 : 127u;
                const int div = max(1, abs(last_b - b + 1));
                uint32_t pen2;
                if constexpr (PEN2_TABLE) pen2 = S.pen2[min(div, 255)];
                else pen2 = (uint32_t)max(a.p1, a.p2 / div);
                const uint32_t v = min(min(L, bm), min(bp, m + pen2));
                const uint32_t l = c + (v - m);
                const uint32_t mn = group_min<LR>(l, lane);
                o = valid ? l : 0u;
                L = valid ? l : L;
                m = valid ? mn : 127u;
                last_b = valid ? b : last_b;
            }
            S.out[(sub * K + s) * LR + t] = (uint8_t)o;
        }
        __builtin_amdgcn_wave_barrier();
        {   // this block's results, one window per lane
            const int step = blk * K + lstep;
            if (lline < nlines && step < n) {
                uint4* dst = reinterpret_cast<uint4*>(out + pix_of(lline, step) * LR);
                const uint4* src = reinterpret_cast<const uint4*>(S.out + lane * LR);
#pragma unroll
                for (int q = 0; q < NV; ++q) dst[q] = src[q];
            }
        }
        if (blk + 1 < nblk) {
            stage();
            if (blk + 2 < nblk) fetch(blk + 2);
        }
    }
}

template <int LR>
__global__ __launch_bounds__(64) void agg_init_kernel(PyrLevel a) {
    __shared__ __align__(16) AggInitLds<LR> lds;
    constexpr int SPW = 64 / LR;
    const int nh = (a.h + SPW - 1) / SPW, nv = (a.w + SPW - 1) / SPW;
    const int item = blockIdx.x;
    int dir, group;
    if (item < nh) { dir = 1; group = item; }
    else if (item < 2 * nh) { dir = 3; group = item - nh; }
    else if (item < 2 * nh + nv) { dir = 0; group = item - 2 * nh; }
    else { dir = 2; group = item - 2 * nh - nv; }
    if (a.p2 <= 254) agg_init_lines<LR, true>(a, dir, group, threadIdx.x, lds);
    else agg_init_lines<LR, false>(a, dir, group, threadIdx.x, lds);
}

// The same aggregation with two disparities per lane as packed u16 pairs:
// LR / 2 lanes per scanline, SPW = 128 / LR scanlines per wave, and every
// VALU instruction of the step works on two cells.  The window shift goes
// through LDS.  Each step writes the scanline's L + P1 into a row of
// overlapping pairs, C[e] = (L'[e - PAD], L'[e - PAD + 1]) + P1, padded on
// both sides with 127.  A lane then reads the four neighbour terms its two
// disparities need, L'[2k + b - b' - 1 .. 2k + b - b' + 2] + P1, as the two
// dwords C[e0] and C[e0 + 2]: dword-aligned whatever the parity of b - b'.
// (Reading the four u16 of a plain row is a misaligned 8-byte LDS access,
// and its stalls took most of the kernel's time.)  A pad is exactly the 127
// of a neighbour outside the window, so no range test remains.  b - b' is
// clamped to +-(LR + 1), past which every term is a pad, and, separately, to
// [-254, 256] to index a table of max(P1, P2 / max(1, |b' - b + 1|)).
//
// Exact when P1 <= 127 and P2 <= 254: the true (unbounded) values of the
// recurrence then lie in [-382, 763] (l <= c + P2 <= 509, and l >= 127 - m
// after a neighbour outside the window), so the minima of their u16
// representatives pick the same entries as those of their u32
// representatives, and the stored low bytes agree.  Other penalties take
// agg_init_kernel.
typedef unsigned short pk16 __attribute__((ext_vector_type(2)));
typedef unsigned int pyr_u32x4 __attribute__((ext_vector_type(4)));

// Block stores of the window aggregation: one LR-byte window per pixel, as
// buffer stores from a base that moves by a fixed step per block, each pixel
// of a block at an offset from it fixed for the whole walk.  A pixel past the
// scanlines or the line's end gets an offset beyond the resource's range and
// its store is dropped, so every block issues the same stores with no branch
// around them: the compiler then counts the prefetch ring's outstanding loads
// exactly instead of draining every load and store at each block.  Block blk
// of scanline group l0 holds pixel (scanline l0 + ln, step blk K + j); the
// base is the lowest such pixel, offsets stay below (64 + K) W LR bytes.
template <int LR, int K>
struct BlockStores {
    int64_t base0, step;   // base pixel at block 0, and its move per block
    __device__ BlockStores(bool horiz, bool fwd, int W, int n, int l0) {
        base0 = horiz ? (int64_t)l0 * W + (fwd ? 0 : n - K) : (fwd ? 0 : (int64_t)(n - K) * W) + l0;
        step = (horiz ? 1 : (int64_t)W) * (fwd ? K : -K);
    }
    // byte offset of (ln, j) from the base
    __device__ static uint32_t offset(bool horiz, bool fwd, int W, int ln, int j) {
        const int jj = fwd ? j : K - 1 - j;
        return (uint32_t)((horiz ? ln * W + jj : jj * W + ln) * LR);
    }
    __device__ void store(uint8_t* out, int blk, uint32_t off, bool ok, const uint8_t* src) const {
        // in the reverse directions the base of a line's last, partial block
        // (and of the blocks past it) lies before the volume: only pixels
        // inside the line are in range, and their offsets land inside it
        const uintptr_t base = (uintptr_t)out + (uintptr_t)((base0 + blk * step) * LR);
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, 0x7fffffff, 0x00020000);
        const uint32_t o = ok ? off : 0x80000000u;
#pragma unroll
        for (int v = 0; v < LR / 16; ++v) {
            const uint4 w = reinterpret_cast<const uint4*>(src)[v];
            __builtin_amdgcn_raw_buffer_store_b128(pyr_u32x4{w.x, w.y, w.z, w.w}, rs, (int)(o + 16 * v), 0, 0);
        }
    }
};

constexpr int kPkDeltaLo = -254, kPkDeltaHi = 256;   // table range of b - b'
// Blocks of cost loads in flight per wave.  A block is LR / 2 steps; its
// loads are issued kPkAhead blocks before they are staged, so the HBM
// latency (1-2 us under load) is spread over that many blocks of steps.
constexpr int kPkAhead = 4;

template <int LR>
struct AggInitPkLds {
    static constexpr int SPW = 128 / LR;
    static constexpr int PAD = LR + 2;               // even: the lanes' pair stores stay 8-byte aligned
    static constexpr int ROW = PAD + 2 * LR + 2;     // even; reads reach PAD + 2 LR
    uint32_t cost[SPW][LR / 2][LR / 2];      // [line][pair k][step]: the pair's costs as u16
    uint8_t out[64 * LR];                    // [line][step][t]
    int16_t init[64];                        // [line][step]
    uint32_t pen2[kPkDeltaHi - kPkDeltaLo + 1];   // by b - b', both halves
    uint32_t win[SPW][ROW];                  // overlapping pairs, see above
};

__device__ __forceinline__ uint32_t pk_bits(pk16 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ pk16 pk_of(uint32_t v) { return __builtin_bit_cast(pk16, v); }

// v_perm selector of bytes 2 (q + 1) .. 2 (q + 1) + 3 of {hi, lo} for a
// window shift q in {-1, 0, 1}: the dword starting one disparity below, at,
// or one above the middle of the pair boundary.  Other q (lanes whose step
// is skipped) are clamped into that range; their result is not used.
// Three full-rate instructions: q clamped by v_med3, then 24-bit multiply-adds
// (written out: the compiler merges the steps into one quarter-rate 32-bit
// multiply).  The constants come in registers (no literals in VOP3).
__device__ __forceinline__ uint32_t shift_sel(int q, uint32_t k0202, uint32_t k10001, uint32_t kbase) {
    uint32_t t, u, sel;
    asm("v_med3_i32 %0, %3, -1, 1\n\t"
        "v_mad_i32_i24 %1, %0, %4, %4\n\t"     // (q + 1) * 0x0202
        "v_mad_u32_u24 %2, %1, %5, %6"           // u * 0x10001 + 0x03020100
        : "=&v"(t), "=&v"(u), "=v"(sel)
        : "v"(q), "s"(k0202), "v"(k10001), "s"(kbase));
    return sel;
}

// minimum over the LR / 2 lanes of a scanline (STRIDE = 32 / LR apart in a
// DPP row) of both halves; the result holds it in both halves
template <int LR>
__device__ __forceinline__ pk16 pk_group_min(pk16 v) {
    constexpr int STRIDE = 32 / LR;
    v = __builtin_elementwise_min(v, v.yx);
    uint32_t m = pk_bits(v);                 // equal halves: u32 order = u16 order
    m = min(m, dppu<0x120 + STRIDE>(m));     // row_ror by STRIDE, 2 STRIDE, ...
    m = min(m, dppu<0x120 + 2 * STRIDE>(m));
    m = min(m, dppu<0x120 + 4 * STRIDE>(m));
    if constexpr (8 * STRIDE < 16) m = min(m, dppu<0x120 + 8 * STRIDE>(m));
    return pk_of(m);
}

template <int LR>
__device__ __forceinline__ void agg_init_pk_lines(const PyrLevel& a, int dir, int group, int lane,
                                                  AggInitPkLds<LR>& S) {
    constexpr int LPL = LR / 2, SPW = 128 / LR, K = LPL;
    constexpr int PAD = AggInitPkLds<LR>::PAD, ROW = AggInitPkLds<LR>::ROW;
    constexpr int NV = LR / 16;
    constexpr uint32_t kPadPair = 127u * 0x10001u;
    // the lanes of a scanline sit STRIDE = 16 / LPL apart in one DPP row (see
    // agg_init_pk2_lines): the edge pads stay in unwritten lanes
    constexpr int STRIDE = 16 / LPL;
    const int sub = (lane / 16) * STRIDE + lane % STRIDE;   // recurrence: line
    const int k = (lane % 16) / STRIDE;                     // disparity pair
    const int lsub = lane / K, lstep = lane % K;     // loads / stores: line, step
    const bool horiz = (dir == 1 || dir == 3);
    const bool fwd = (dir == 0 || dir == 1);
    const int W = a.w, H = a.h;
    const int nlines = horiz ? H : W;
    const int n = horiz ? W : H;
    auto pix_of = [&](int line, int step) -> int64_t {
        line = min(line, nlines - 1);
        step = min(step, n - 1);
        const int pos = fwd ? step : n - 1 - step;
        return horiz ? (int64_t)line * W + pos : (int64_t)pos * W + line;
    };
    const int line = min(group * SPW + sub, nlines - 1);
    const int lline = group * SPW + lsub;
    uint8_t* out = (dir == 0) ? a.path[0] : (dir == 1) ? a.path[1] : (dir == 2) ? a.path[2] : a.path[3];

    // a ring of kPkAhead blocks of loads in flight (slot = block mod kPkAhead)
    uint4 ncs[kPkAhead][NV];
    int nbs[kPkAhead];
    auto fetch = [&](int slot, int blk) __attribute__((always_inline)) {
        const int64_t p = pix_of(lline, blk * K + lstep);
        const uint4* src = reinterpret_cast<const uint4*>(a.cost + p * LR);
#pragma unroll
        for (int v = 0; v < NV; ++v) ncs[slot][v] = src[v];
        nbs[slot] = a.init[p];
    };
    // the block in LDS, transposed so that a lane later reads the cost pairs
    // of all K steps of its own (line, pair) with wide loads: pair j of
    // this lane's pixel goes to cost[line][j][step], zero-extended to u16
    auto stage = [&](int slot) __attribute__((always_inline)) {
        uint32_t wv[LR / 4];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            wv[4 * v] = ncs[slot][v].x; wv[4 * v + 1] = ncs[slot][v].y;
            wv[4 * v + 2] = ncs[slot][v].z; wv[4 * v + 3] = ncs[slot][v].w;
        }
#pragma unroll
        for (int j = 0; j < LR / 2; ++j)
            S.cost[lsub][j][lstep] = __builtin_amdgcn_perm(0u, wv[j / 2], (j & 1) ? 0x0c030c02u : 0x0c010c00u);
        S.init[lane] = (int16_t)nbs[slot];
    };

    // pads, and the P2 table by offset q = b - b' (divisor |1 - q|)
    for (int e = lane; e < SPW * ROW; e += 64) (&S.win[0][0])[e] = kPadPair;
    for (int e = lane; e <= kPkDeltaHi - kPkDeltaLo; e += 64) {
        const int q = e + kPkDeltaLo;
        const uint32_t v = (uint32_t)max(a.p1, a.p2 / max(1, abs(1 - q)));
        S.pen2[e] = v | (v << 16);
    }
    __builtin_amdgcn_wave_barrier();

    uint32_t* wrow = &S.win[sub][0];
    const uint32_t* wlane = wrow + (PAD - 1 + 2 * k);     // + q: the pair before this lane's first term
    const bool last_k = k == LPL - 1;
    uint8_t* obase = S.out + sub * K * LR + 2 * k;
    const pk16 p1v = pk_of((uint32_t)a.p1 * 0x10001u);
    const pk16 m127 = pk_of(kPadPair);
    const uint32_t k10001 = 0x10001u;
    // clamp bounds kept in registers (one v_med3 each per step) instead of
    // rematerialised each step
    int qlo = kPkDeltaLo, qhi = kPkDeltaHi, wlo = -(LR + 1);
    asm volatile("" : "+v"(qlo), "+v"(qhi), "+v"(wlo));
    // P2 terms of the offsets -1, 0 and 1 (divisors 2, 1 and 1)
    const pk16 pen2_m1 = pk_of(S.pen2[-1 - kPkDeltaLo]);
    const pk16 pen2_01 = pk_of(S.pen2[0 - kPkDeltaLo]);
    // this lane's pairs of L + P1 into the window row: C[PAD + 2k] (its own
    // pair), C[PAD + 2k + 1] (its high half and the next lane's low half or a
    // pad) and, from lane 0, C[PAD - 1] (a pad and its low half)
    auto put = [&](uint32_t own) __attribute__((always_inline)) {
        uint32_t nxt = dppu<0x100 + STRIDE>(own);         // row_shl: lane k + 1
        nxt = last_k ? kPadPair : nxt;
        const uint32_t mid = __builtin_amdgcn_alignbit(nxt, own, 16);
        *reinterpret_cast<uint2*>(wrow + PAD + 2 * k) = make_uint2(own, mid);
        if (k == 0) wrow[PAD - 1] = (own << 16) | 127u;
    };

    pk16 L = pk_of(0u), m = pk_of(0u);
    uint32_t nxt_r = kPadPair, prv_r = kPadPair;   // pads kept in the last / first lane of a line
    int last_b = 0;
    // A11: vertical directions seed b' from row 0 whichever end they start at
    const int seed_b = horiz ? 0 : a.init[line];
    const int nblk = (n + K - 1) / K;
    const BlockStores<LR, K> stores(horiz, fwd, W, n, group * SPW);
    const uint32_t woff = BlockStores<LR, K>::offset(horiz, fwd, W, lsub, lstep);
    const bool wline = lline < nlines;
    auto process = [&](int blk) __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        // the bases of the block's K steps of this lane's line and this
        // lane's cost pairs of those steps, in registers: no LDS read on the
        // step's chain
        uint32_t bw[K / 2], cwv[K];
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(&S.init[sub * K])[i];
            bw[4 * i] = v.x; bw[4 * i + 1] = v.y; bw[4 * i + 2] = v.z; bw[4 * i + 3] = v.w;
        }
#pragma unroll
        for (int i = 0; i < K / 4; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(&S.cost[sub][k][0])[i];
            cwv[4 * i] = v.x; cwv[4 * i + 1] = v.y; cwv[4 * i + 2] = v.z; cwv[4 * i + 3] = v.w;
        }
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int b = (int)(int16_t)(bw[s / 2] >> (16 * (s & 1)));
            const pk16 c = pk_of(cwv[s]);
            pk16 o;
            if (s == 0 && blk == 0) {
                L = c;
                m = pk_group_min<LR>(L);
                last_b = horiz ? b : seed_b;
                o = L;
            } else {
                const bool valid = b != kInvalidShort;
                const int q = b - last_b;
                const uint32_t own = pk_bits(L + p1v);
                pk16 bm, bp, pen2;                        // neighbour terms, + P1 already
                if (__builtin_expect(__builtin_amdgcn_ballot_w64(valid && (q < -1 || q > 1)) == 0, 1)) {
                    // every line shifts by at most one disparity: the terms are
                    // this lane's pair and its neighbours' (DPP), no LDS
                    // lane k + 1 and lane k - 1 (row_shl / row_shr by STRIDE,
                    // out-of-row sources not written: the pads stay)
                    nxt_r = (uint32_t)__builtin_amdgcn_update_dpp((int)nxt_r, (int)own, 0x100 + STRIDE, 0xF, 0xF,
                                                                  false);
                    prv_r = (uint32_t)__builtin_amdgcn_update_dpp((int)prv_r, (int)own, 0x110 + STRIDE, 0xF, 0xF,
                                                                  false);
                    const uint32_t nxt = nxt_r, prv = prv_r;
                    // the pair starting at disparity 2k + q - 1 (2k + q + 1) is
                    // bytes 2 (q + 1) .. 2 (q + 1) + 3 of {own, prv} ({nxt, own})
                    const uint32_t sel = shift_sel(q, 0x0202u, k10001, 0x03020100u);
                    bm = pk_of(__builtin_amdgcn_perm(own, prv, sel));
                    bp = pk_of(__builtin_amdgcn_perm(nxt, own, sel));
                    const uint32_t mneg = 0u - (uint32_t)(q < 0);
                    pen2 = pk_of((pk_bits(pen2_m1) & mneg) | (pk_bits(pen2_01) & ~mneg));
                } else {
                    put(own);      // the window row holds L only when a step needs it
                    const int qt = min(max(q, qlo), qhi);
                    const int qw = min(max(q, wlo), LR + 1);
                    uint32_t xm = wlane[qw], xp = wlane[qw + 2], pb = S.pen2[qt - kPkDeltaLo];
                    // complete these reads inside the branch, so that no LDS read
                    // is pending where the two paths meet (the fast path would
                    // otherwise wait for every LDS operation, the output stores
                    // included)
                    asm volatile("" : "+v"(xm), "+v"(xp), "+v"(pb));
                    bm = pk_of(xm);
                    bp = pk_of(xp);
                    pen2 = pk_of(pb);
                }
                const pk16 v = __builtin_elementwise_min(__builtin_elementwise_min(L, bm),
                                                         __builtin_elementwise_min(bp, m + pen2));
                const pk16 l = c + (v - m);
                const pk16 mn = pk_group_min<LR>(l);
                // selects as bit operations (v_bfi), so the step stays one block
                const uint32_t vm = 0u - (uint32_t)valid;
                o = pk_of(pk_bits(l) & vm);
                L = pk_of((pk_bits(l) & vm) | (pk_bits(L) & ~vm));
                m = pk_of((pk_bits(mn) & vm) | (pk_bits(m127) & ~vm));
                last_b = (int)(((uint32_t)b & vm) | ((uint32_t)last_b & ~vm));
            }
            // low bytes of both halves
            *reinterpret_cast<uint16_t*>(obase + s * LR) =
                (uint16_t)__builtin_amdgcn_perm(0u, pk_bits(o), 0x0c0c0200u);
        }
        __builtin_amdgcn_wave_barrier();
        // this block's results, one window per lane
        stores.store(out, blk, woff, wline && lstep < n - blk * K, S.out + lane * LR);
    };
    // the prologue issues the ring's slots in the loop's order: the wait
    // counts at the loop head merge both paths
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kPkAhead; ++u) {
        fetch(u, u);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int blk0 = 0; blk0 < nblk; blk0 += kPkAhead) {
#pragma unroll
        for (int u = 0; u < kPkAhead; ++u) {
            // blocks past the last one (at most kPkAhead - 1) walk the
            // line's last pixel again; their stores are dropped
            const int blk = blk0 + u;
            __builtin_amdgcn_wave_barrier();
            stage(u);
            fetch(u, blk + kPkAhead);
            process(blk);
        }
    }
}

template <int LR>
__global__ __launch_bounds__(64) void agg_init_pk_kernel(PyrLevel a) {
    __shared__ __align__(16) AggInitPkLds<LR> lds;
    constexpr int SPW = 128 / LR;
    const int nh = (a.h + SPW - 1) / SPW, nv = (a.w + SPW - 1) / SPW;
    const int item = blockIdx.x;
    int dir, group;
    if (item < nh) { dir = 1; group = item; }
    else if (item < 2 * nh) { dir = 3; group = item - nh; }
    else if (item < 2 * nh + nv) { dir = 0; group = item - 2 * nh; }
    else { dir = 2; group = item - 2 * nh - nv; }
    agg_init_pk_lines<LR>(a, dir, group, threadIdx.x, lds);
}

// The packed aggregation with four disparities per lane (two u16 pairs, A =
// 4k, 4k+1 and B = 4k+2, 4k+3): LR / 4 lanes per scanline, 256 / LR
// scanlines per wave.  Per step a lane needs the neighbour pairs starting
// at 4k + q - 1, 4k + q + 1 and 4k + q + 3 (q = b - b'), the middle one
// serving both of its pairs, so a step costs about as many instructions as
// the two-disparity form for twice the cells.  It has half as many waves,
// which suits the largest level; the launch picks the form.  Same window
// row (overlapping pairs, 127 pads), table and exactness range.
template <int LR>
struct AggInitPk2Lds {
    static constexpr int LPL = LR / 4, SPW = 64 / LPL, K = LR / 2;
    static constexpr int PAD = (LR + 2 + 3) / 4 * 4;         // a multiple of 4: 16-byte row stores
    static constexpr int ROW = (PAD + 2 * LR + 1 + 3) / 4 * 4;
    uint32_t cost[SPW][LR / 2][K];           // [line][pair][step]: the pair's costs as u16
    uint8_t out[2 * 64 * LR];                // [line][step][t]
    int16_t init[2 * 64];                    // [line][step]
    uint32_t pen2[kPkDeltaHi - kPkDeltaLo + 1];
    uint32_t win[SPW][ROW];
};

template <int LR>
__device__ __forceinline__ void agg_init_pk2_lines(const PyrLevel& a, int dir, int group, int lane,
                                                   AggInitPk2Lds<LR>& S) {
    using Lds = AggInitPk2Lds<LR>;
    constexpr int LPL = Lds::LPL, SPW = Lds::SPW, K = Lds::K, PAD = Lds::PAD, ROW = Lds::ROW;
    constexpr int NV = LR / 16;
    constexpr uint32_t kPadPair = 127u * 0x10001u;
    static_assert(SPW * K == 128, "a block is two pixels per lane");
    // The lanes of a scanline sit STRIDE = 16 / LPL apart in one DPP row, so a
    // neighbouring quad is a row shift by STRIDE.  The line's first and last
    // lanes then read from outside the row, which leaves a move's old value in
    // place: the 127 pads stay in those lanes of two persistent registers,
    // with no select per step.
    constexpr int STRIDE = 16 / LPL;
    const int sub = (lane / 16) * STRIDE + lane % STRIDE;   // recurrence: line
    const int k = (lane % 16) / STRIDE;                     // disparity quad
    const bool horiz = (dir == 1 || dir == 3);
    const bool fwd = (dir == 0 || dir == 1);
    const int W = a.w, H = a.h;
    const int nlines = horiz ? H : W;
    const int n = horiz ? W : H;
    auto pix_of = [&](int line, int step) -> int64_t {
        line = min(line, nlines - 1);
        step = min(step, n - 1);
        const int pos = fwd ? step : n - 1 - step;
        return horiz ? (int64_t)line * W + pos : (int64_t)pos * W + line;
    };
    const int line = min(group * SPW + sub, nlines - 1);
    uint8_t* out = (dir == 0) ? a.path[0] : (dir == 1) ? a.path[1] : (dir == 2) ? a.path[2] : a.path[3];

    // loads and stores: lane j handles the block's pixels j and j + 64 in
    // [line][step] order
    // a ring of kPkAhead blocks of loads in flight (slot = block mod kPkAhead)
    uint4 nc[kPkAhead][2][NV];
    int nb[kPkAhead][2];
    auto fetch = [&](int slot, int blk) __attribute__((always_inline)) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = lane + 64 * h;
            const int64_t p = pix_of(group * SPW + idx / K, blk * K + idx % K);
            const uint4* src = reinterpret_cast<const uint4*>(a.cost + p * LR);
#pragma unroll
            for (int v = 0; v < NV; ++v) nc[slot][h][v] = src[v];
            nb[slot][h] = a.init[p];
        }
    };
    auto stage = [&](int slot) __attribute__((always_inline)) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int idx = lane + 64 * h;
            uint32_t wv[LR / 4];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                wv[4 * v] = nc[slot][h][v].x; wv[4 * v + 1] = nc[slot][h][v].y;
                wv[4 * v + 2] = nc[slot][h][v].z; wv[4 * v + 3] = nc[slot][h][v].w;
            }
#pragma unroll
            for (int j = 0; j < LR / 2; ++j)
                S.cost[idx / K][j][idx % K] =
                    __builtin_amdgcn_perm(0u, wv[j / 2], (j & 1) ? 0x0c030c02u : 0x0c010c00u);
            S.init[idx] = (int16_t)nb[slot][h];
        }
    };

    for (int e = lane; e < SPW * ROW; e += 64) (&S.win[0][0])[e] = kPadPair;
    for (int e = lane; e <= kPkDeltaHi - kPkDeltaLo; e += 64) {
        const int q = e + kPkDeltaLo;
        const uint32_t v = (uint32_t)max(a.p1, a.p2 / max(1, abs(1 - q)));
        S.pen2[e] = v | (v << 16);
    }
    __builtin_amdgcn_wave_barrier();

    uint32_t* wrow = &S.win[sub][0];
    const uint32_t* wlane = wrow + (PAD - 1 + 4 * k);     // + q: the first neighbour pair
    const bool last_k = k == LPL - 1;
    uint8_t* obase = S.out + sub * K * LR + 4 * k;
    const pk16 p1v = pk_of((uint32_t)a.p1 * 0x10001u);
    const pk16 m127 = pk_of(kPadPair);
    const uint32_t k10001 = 0x10001u;
    const pk16 pen2_m1 = pk_of(S.pen2[-1 - kPkDeltaLo]);
    const pk16 pen2_01 = pk_of(S.pen2[0 - kPkDeltaLo]);
    int qlo = kPkDeltaLo, qhi = kPkDeltaHi, wlo = -(LR + 1);
    asm volatile("" : "+v"(qlo), "+v"(qhi), "+v"(wlo));
    // the window row from this lane's pairs of L + P1 (A, B): C[PAD + 4k ..
    // PAD + 4k + 3] in one 16-byte store, and C[PAD - 1] from lane 0
    auto put = [&](uint32_t oa, uint32_t ob) __attribute__((always_inline)) {
        uint32_t nxt = dppu<0x100 + STRIDE>(oa);          // row_shl: lane k + 1's A
        nxt = last_k ? kPadPair : nxt;
        const uint4 v = make_uint4(oa, __builtin_amdgcn_alignbit(ob, oa, 16), ob,
                                   __builtin_amdgcn_alignbit(nxt, ob, 16));
        *reinterpret_cast<uint4*>(wrow + PAD + 4 * k) = v;
        if (k == 0) wrow[PAD - 1] = (oa << 16) | 127u;
    };
    auto group_min4 = [&](pk16 v) __attribute__((always_inline)) {
        v = __builtin_elementwise_min(v, v.yx);
        uint32_t m = pk_bits(v);
        m = min(m, dppu<0x120 + STRIDE>(m));              // row_ror by STRIDE, 2 STRIDE, ...
        if constexpr (2 * STRIDE < 16) m = min(m, dppu<0x120 + 2 * STRIDE>(m));
        if constexpr (4 * STRIDE < 16) m = min(m, dppu<0x120 + 4 * STRIDE>(m));
        return pk_of(m);
    };

    pk16 LA = pk_of(0u), LB = pk_of(0u), m = pk_of(0u);
    int last_b = 0;
    // A11: vertical directions seed b' from row 0 whichever end they start at
    const int seed_b = horiz ? 0 : a.init[line];
    const int nblk = (n + K - 1) / K;
    const BlockStores<LR, K> stores(horiz, fwd, W, n, group * SPW);
    uint32_t woff[2];
    bool wline[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int idx = lane + 64 * h;
        woff[h] = BlockStores<LR, K>::offset(horiz, fwd, W, idx / K, idx % K);
        wline[h] = group * SPW + idx / K < nlines;
    }
    uint32_t nxt_r = kPadPair, prv_r = kPadPair;   // pads kept in the last / first lane of a line
    // neighbour pairs at 4k + q - 1, 4k + q + 1, 4k + q + 3 for a base step q
    // in {-1, 0, 1}: bytes 2 (q + 1) .. 2 (q + 1) + 3 of {oa, prv}, {ob, oa},
    // {nxt, ob}, with `sel` the selector of q
    auto dpp_terms = [&](uint32_t oa, uint32_t ob, uint32_t sel, int q, uint32_t& x0, uint32_t& x1,
                         uint32_t& x2, pk16& pen2) __attribute__((always_inline)) {
        // lane k + 1's A and lane k - 1's B (row_shl / row_shr by STRIDE,
        // out-of-row sources not written)
        nxt_r = (uint32_t)__builtin_amdgcn_update_dpp((int)nxt_r, (int)oa, 0x100 + STRIDE, 0xF, 0xF, false);
        prv_r = (uint32_t)__builtin_amdgcn_update_dpp((int)prv_r, (int)ob, 0x110 + STRIDE, 0xF, 0xF, false);
        x0 = __builtin_amdgcn_perm(oa, prv_r, sel);
        x1 = __builtin_amdgcn_perm(ob, oa, sel);
        x2 = __builtin_amdgcn_perm(nxt_r, ob, sel);
        const uint32_t mneg = 0u - (uint32_t)(q < 0);
        pen2 = pk_of((pk_bits(pen2_m1) & mneg) | (pk_bits(pen2_01) & ~mneg));
    };
    // this step's L of the lane's two pairs from the previous L, m and the
    // neighbour terms
    auto relax = [&](pk16 cA, pk16 cB, uint32_t x0, uint32_t x1, uint32_t x2, pk16 pen2, pk16& lA,
                     pk16& lB) __attribute__((always_inline)) {
        const pk16 t = __builtin_elementwise_min(pk_of(x1), m + pen2);
        const pk16 vA = __builtin_elementwise_min(__builtin_elementwise_min(LA, pk_of(x0)), t);
        const pk16 vB = __builtin_elementwise_min(__builtin_elementwise_min(LB, pk_of(x2)), t);
        lA = cA + (vA - m);
        lB = cB + (vB - m);
    };
    auto process = [&](int blk) __attribute__((always_inline)) {
        __builtin_amdgcn_wave_barrier();
        uint32_t bw[K / 2], ca[K], cb[K];
#pragma unroll
        for (int i = 0; i < K / 8; ++i) {
            const uint4 v = reinterpret_cast<const uint4*>(&S.init[sub * K])[i];
            bw[4 * i] = v.x; bw[4 * i + 1] = v.y; bw[4 * i + 2] = v.z; bw[4 * i + 3] = v.w;
        }
#pragma unroll
        for (int i = 0; i < K / 4; ++i) {
            const uint4 va = reinterpret_cast<const uint4*>(&S.cost[sub][2 * k][0])[i];
            const uint4 vb = reinterpret_cast<const uint4*>(&S.cost[sub][2 * k + 1][0])[i];
            ca[4 * i] = va.x; ca[4 * i + 1] = va.y; ca[4 * i + 2] = va.z; ca[4 * i + 3] = va.w;
            cb[4 * i] = vb.x; cb[4 * i + 1] = vb.y; cb[4 * i + 2] = vb.z; cb[4 * i + 3] = vb.w;
        }
#pragma unroll
        for (int s = 0; s < K; ++s) {
            const int b = (int)(int16_t)(bw[s / 2] >> (16 * (s & 1)));
            const pk16 cA = pk_of(ca[s]), cB = pk_of(cb[s]);
            uint32_t ow;   // the low bytes of the four halves
            if (s == 0 && blk == 0) {
                LA = cA;
                LB = cB;
                m = group_min4(__builtin_elementwise_min(LA, LB));
                last_b = horiz ? b : seed_b;
                ow = __builtin_amdgcn_perm(pk_bits(LB), pk_bits(LA), 0x06040200u);
            } else {
                const bool valid = b != kInvalidShort;
                const int q = b - last_b;
                const uint32_t oa = pk_bits(LA + p1v), ob = pk_bits(LB + p1v);
                uint32_t x0, x1, x2;                     // pairs at 4k + q - 1, + 1, + 3
                pk16 pen2;
                if (__builtin_expect(__builtin_amdgcn_ballot_w64(valid && (q < -1 || q > 1)) == 0, 1)) {
                    dpp_terms(oa, ob, shift_sel(q, 0x0202u, k10001, 0x03020100u), q, x0, x1, x2, pen2);
                } else {
                    put(oa, ob);   // the window row holds L only when a step needs it
                    const int qt = min(max(q, qlo), qhi);
                    const int qw = min(max(q, wlo), LR + 1);
                    x0 = wlane[qw];
                    x1 = wlane[qw + 2];
                    x2 = wlane[qw + 4];
                    uint32_t pb = S.pen2[qt - kPkDeltaLo];
                    // complete these reads inside the branch (see agg_init_pk_lines)
                    asm volatile("" : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(pb));
                    pen2 = pk_of(pb);
                }
                pk16 lA, lB;
                relax(cA, cB, x0, x1, x2, pen2, lA, lB);
                const pk16 mn = group_min4(__builtin_elementwise_min(lA, lB));
                const uint32_t vm = 0u - (uint32_t)valid;
                ow = __builtin_amdgcn_perm(pk_bits(lB), pk_bits(lA), 0x06040200u) & vm;
                LA = pk_of((pk_bits(lA) & vm) | (pk_bits(LA) & ~vm));
                LB = pk_of((pk_bits(lB) & vm) | (pk_bits(LB) & ~vm));
                m = pk_of((pk_bits(mn) & vm) | (pk_bits(m127) & ~vm));
                last_b = (int)(((uint32_t)b & vm) | ((uint32_t)last_b & ~vm));
            }
            *reinterpret_cast<uint32_t*>(obase + s * LR) = ow;
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int h = 0; h < 2; ++h) {   // this block's results, one window per pixel
            const int idx = lane + 64 * h;
            stores.store(out, blk, woff[h], wline[h] && idx % K < n - blk * K, S.out + idx * LR);
        }
    };
    // the prologue issues the ring's slots in the loop's order: the wait
    // counts at the loop head merge both paths
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kPkAhead; ++u) {
        fetch(u, u);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int blk0 = 0; blk0 < nblk; blk0 += kPkAhead) {
#pragma unroll
        for (int u = 0; u < kPkAhead; ++u) {
            // blocks past the last one (at most kPkAhead - 1) walk the
            // line's last pixel again; their stores are dropped
            const int blk = blk0 + u;
            __builtin_amdgcn_wave_barrier();
            stage(u);
            fetch(u, blk + kPkAhead);
            process(blk);
        }
    }
}

template <int LR>
__global__ __launch_bounds__(64) void agg_init_pk2_kernel(PyrLevel a) {
    __shared__ __align__(16) AggInitPk2Lds<LR> lds;
    constexpr int SPW = AggInitPk2Lds<LR>::SPW;
    const int nh = (a.h + SPW - 1) / SPW, nv = (a.w + SPW - 1) / SPW;
    const int item = blockIdx.x;
    int dir, group;
    if (item < nh) { dir = 1; group = item; }
    else if (item < 2 * nh) { dir = 3; group = item - nh; }
    else if (item < 2 * nh + nv) { dir = 0; group = item - 2 * nh; }
    else { dir = 2; group = item - 2 * nh - nv; }
    agg_init_pk2_lines<LR>(a, dir, group, threadIdx.x, lds);
}

// WTA with initial disparity (cusgm_ca.cu:3425-3512), one lane per pixel
template <int LR>
__global__ __launch_bounds__(256) void wta_init_kernel(PyrLevel a) {
    const int64_t n = (int64_t)a.w * a.h;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int y = (int)(i / a.w), x = (int)(i - (int64_t)y * a.w);
    const int W = a.w, H = a.h;
    float out = kInvalidDisp;
    const int b = a.init[i];
    if (y < 4 * (H / 4) && !(x > W - 5 || y > H - 5 || x < 5 || y < 5) && b != kInvalidShort) {
        uint32_t S[LR];
#pragma unroll
        for (int t = 0; t < LR; ++t) S[t] = 0u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const uint8_t* pr = (r == 0) ? a.path[0] : (r == 1) ? a.path[1] : (r == 2) ? a.path[2] : a.path[3];
            const uint4* src = reinterpret_cast<const uint4*>(pr + i * LR);
#pragma unroll
            for (int q = 0; q < LR / 16; ++q) {
                const uint4 v = src[q];
                const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int k = 0; k < 16; ++k) S[q * 16 + k] += (w4[k >> 2] >> (8 * (k & 3))) & 0xFFu;
            }
        }
        uint32_t c2 = 0xFFFFFFFFu;
        int ds = 0;
#pragma unroll
        for (int t = 0; t < LR; ++t)
            if (S[t] < c2) { c2 = S[t]; ds = t; }
        if (ds > 1 && ds < LR - 1) {
            uint32_t s1 = 0, s3 = 0;
#pragma unroll
            for (int t = 0; t < LR; ++t) {
                s1 = (t == ds - 1) ? S[t] : s1;
                s3 = (t == ds + 1) ? S[t] : s3;
            }
            const float c1 = (float)s1, c2f = (float)c2, c3 = (float)s3;
            const float aa = c1 - c2f, bb = c3 - c2f;
            const float uniq = (aa < bb) ? aa : bb;
            if (!(uniq <= c2f * (1.0f - a.uniqueness))) {
                const float tt = c1 + c3 - 2.0f * c2f;
                const float den = (1.0f < tt) ? tt : 1.0f;
                const float bias = (float)((double)(c1 - c3 + den) / (double)(2.0f * den));
                out = (float)(ds + b) + bias;
            }
        }
    }
    a.disp[i] = out;
}

// cusgm_df.cu:188-218: INVALID interior pixels take their window maximum
__global__ __launch_bounds__(256) void dilation_kernel(const float* __restrict__ in,
                                                       float* __restrict__ out, int W, int H,
                                                       int wnd) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)W * H) return;
    const int y = (int)(i / W), x = (int)(i - (int64_t)y * W);
    const int r = (wnd - 1) / 2;
    float v = in[i];
    if (v == kInvalidDisp && x > r && x < W - r && y > r && y < H - r)
        for (int a = -r; a <= r; ++a)
            for (int b = -r; b <= r; ++b) v = fmaxf(v, in[(int64_t)(y + b) * W + x + a]);
    out[i] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t pyr_launch_zoom(const PyrZoom& z, hipStream_t s) {
    const int64_t n = 2 * (int64_t)(z.w / 2) * (z.h / 2);
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zoom_kernel, dim3(blocks_for(n)), dim3(256), 0, s, z);
    return hipGetLastError();
}

hipError_t pyr_launch_dilation(const float* in, float* out, int w, int h, int wnd, hipStream_t s) {
    hipLaunchKernelGGL(dilation_kernel, dim3(blocks_for((int64_t)w * h)), dim3(256), 0, s, in, out,
                       w, h, wnd);
    return hipGetLastError();
}

hipError_t pyr_launch_init_next(const PyrInit& a, hipStream_t s) {
    hipLaunchKernelGGL(init_next_kernel, dim3(blocks_for((int64_t)a.w * a.h)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t pyr_launch_cost_init(const PyrLevel& a, hipStream_t s) {
    const dim3 g(blocks_for((int64_t)a.w * a.h)), b(256);
    if (a.level_range == 16) hipLaunchKernelGGL(cost_init_kernel<16>, g, b, 0, s, a);
    else hipLaunchKernelGGL(cost_init_kernel<32>, g, b, 0, s, a);
    return hipGetLastError();
}

// Whether the packed-u16 window aggregation is exact for these penalties.
static bool agg_init_packed_ok(const PyrLevel& a) { return a.p1 <= 127 && a.p2 <= 254; }

static std::atomic<int> g_window_form{SGM_WINDOW_AUTO};
int window_form() { return g_window_form.load(); }
void set_window_form(int form) { g_window_form.store(form); }

// SIMDs of the current device (4 per compute unit), looked up once per device
static int device_simds() {
    static std::mutex mu;
    static int simds[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
    std::lock_guard<std::mutex> lock(mu);
    if (!simds[dev]) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        simds[dev] = 4 * cus;
    }
    return simds[dev];
}

hipError_t pyr_launch_aggregate_init(const PyrLevel& a, hipStream_t s) {
    auto items = [&](int spw) { return 2 * ((a.h + spw - 1) / spw) + 2 * ((a.w + spw - 1) / spw); };
    int form = window_form();
    if (!agg_init_packed_ok(a)) form = SGM_WINDOW_U32;
    if (form == SGM_WINDOW_AUTO)
        form = items(256 / a.level_range) >= device_simds() ? SGM_WINDOW_QUADS : SGM_WINDOW_PAIRS;
    if (form == SGM_WINDOW_QUADS) {
        const int n = items(256 / a.level_range);
        if (a.level_range == 16) hipLaunchKernelGGL(agg_init_pk2_kernel<16>, dim3(n), dim3(64), 0, s, a);
        else hipLaunchKernelGGL(agg_init_pk2_kernel<32>, dim3(n), dim3(64), 0, s, a);
        return hipGetLastError();
    }
    if (form == SGM_WINDOW_PAIRS) {
        const int n = items(128 / a.level_range);
        if (a.level_range == 16) hipLaunchKernelGGL(agg_init_pk_kernel<16>, dim3(n), dim3(64), 0, s, a);
        else hipLaunchKernelGGL(agg_init_pk_kernel<32>, dim3(n), dim3(64), 0, s, a);
        return hipGetLastError();
    }
    if (a.level_range == 16) {
        const int n = items(4);
        hipLaunchKernelGGL(agg_init_kernel<16>, dim3(n), dim3(64), 0, s, a);
    } else {
        const int n = items(2);
        hipLaunchKernelGGL(agg_init_kernel<32>, dim3(n), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t pyr_launch_wta_init(const PyrLevel& a, hipStream_t s) {
    const dim3 g(blocks_for((int64_t)a.w * a.h)), b(256);
    if (a.level_range == 16) hipLaunchKernelGGL(wta_init_kernel<16>, g, b, 0, s, a);
    else hipLaunchKernelGGL(wta_init_kernel<32>, g, b, 0, s, a);
    return hipGetLastError();
}

}  // namespace fsgm
