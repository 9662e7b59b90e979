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

#include "pyr_internal.h"

#pragma clang fp contract(off)

namespace fsgm {
namespace {

constexpr int kInvalidShort = -32768;

__device__ __forceinline__ uint32_t bperm(int src_lane, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src_lane << 2, (int)v);
}

template <int CTRL>
__device__ __forceinline__ uint32_t dppu(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
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
#pragma unroll
    for (int t = 0; t < LR; ++t) {
        const int xr = x - t - b;
        uint32_t c = invalid;
        if (xr > 0 && xr < a.w) c = (uint32_t)__popcll(cl ^ cr[xr]) + cost_adjust(t, LR);
        words[t >> 2] |= (c & 0xFFu) << (8 * (t & 3));
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

hipError_t pyr_launch_aggregate_init(const PyrLevel& a, hipStream_t s) {
    if (a.level_range == 16) {
        const int items = 2 * ((a.h + 3) / 4) + 2 * ((a.w + 3) / 4);
        hipLaunchKernelGGL(agg_init_kernel<16>, dim3(items), dim3(64), 0, s, a);
    } else {
        const int items = 2 * ((a.h + 1) / 2) + 2 * ((a.w + 1) / 2);
        hipLaunchKernelGGL(agg_init_kernel<32>, dim3(items), dim3(64), 0, s, a);
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
