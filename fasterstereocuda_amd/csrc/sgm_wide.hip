// SGM aggregation with the matching cost computed from the census words on
// the fly (no stored cost volume), for D = 64, 128, 256 (gfx950, wave64).
//
// Same arithmetic as the stored-cost kernels in sgm_kernels.hip (the
// reference's cusgm_cc.cu:118-172 cost and cusgm_ca.cu:1480-1888 recurrence,
// WTA cusgm_ca.cu:3606-3809), so every volume and disparity is bit-identical;
// only where the cost comes from and how lanes are laid out differ:
//
//   * one scanline = D/DPL lanes ("a line"), DPL disparities per lane (16
//     unless FSGM_DPL_* lowers it) held as packed u16 pairs; a wave covers
//     64 DPL / D scanlines.  Neighbour terms
//     d-1 / d+1 across lanes come from row_shr:1 / row_shl:1 DPP moves, the
//     line minimum from quad_perm / mirror DPP reductions.
//   * horizontal directions keep each lane's 16 partner census words in a
//     register window: one step slides it by one disparity, so a lane takes
//     the word its left neighbour drops (one DPP move) and the line's first
//     lane loads the entering word.  The window slots rotate at compile time
//     over a 16-step unrolled block.  A direction whose window would slide the
//     other way runs with the disparity order reversed inside the lanes (the
//     recurrence is symmetric in d); only the byte order of its stores changes.
//   * vertical directions load the census words of the next PF steps into
//     registers (PF = FSGM_PF_V or FSGM_PF_F) and stage, per step, the wave's
//     own words and the D + lines - 1 partner words into one of two LDS
//     slots private to the (single-wave) workgroup; each lane reads its DPL
//     partners from there.
//   * cost(x, t) = popcount(own(x) ^ partner(x, t)): 2 v_xor + 2 v_bcnt.
#include <algorithm>

#include "sgm_device.h"
#include "sgm_internal.h"

#pragma clang fp contract(off)

// prefetch depths (steps) of the top->bottom and fused kernels
#ifndef FSGM_PF_V
#define FSGM_PF_V 8
#endif
#ifndef FSGM_PF_F
#define FSGM_PF_F 4
#endif

namespace fsgm {
namespace {

using namespace dev;

constexpr int kQuadXor1 = 0xB1;        // quad_perm [1, 0, 3, 2]
constexpr int kQuadXor2 = 0x4E;        // quad_perm [2, 3, 0, 1]
constexpr int kRowHalfMirror = 0x141;  // lane i <-> 7 - i within 8 lanes
constexpr int kRowMirror = 0x140;      // lane i <-> 15 - i within 16 lanes

// reductions over the LPL lanes of a line (LPL = 4, 8, 16); every lane gets it
template <int LPL>
__device__ __forceinline__ uint32_t line_min(uint32_t v) {
    v = min(v, dpp0<kQuadXor1>(v));
    v = min(v, dpp0<kQuadXor2>(v));
    if constexpr (LPL >= 8) v = min(v, dpp0<kRowHalfMirror>(v));
    if constexpr (LPL >= 16) v = min(v, dpp0<kRowMirror>(v));
    return v;
}
// value of the previous / next lane of the line; the line's first / last
// lane gets `edge`
template <int LPL>
__device__ __forceinline__ uint32_t from_prev(uint32_t v, uint32_t edge, bool first) {
    if constexpr (LPL == 16) {
        return dpp<kRowShr1>(edge, v);
    } else {
        const uint32_t s = dpp0<kRowShr1>(v);   // every lane moves; the first then takes `edge`
        return first ? edge : s;
    }
}
template <int LPL>
__device__ __forceinline__ uint32_t from_next(uint32_t v, uint32_t edge, bool last) {
    if constexpr (LPL == 16) {
        return dpp<kRowShl1>(edge, v);
    } else {
        const uint32_t s = dpp0<kRowShl1>(v);
        return last ? edge : s;
    }
}

__device__ __forceinline__ uint32_t popc64(uint64_t a) { return (uint32_t)__popcll(a); }
// popcount(a) + acc as two accumulating v_bcnt_u32_b32.  Written out, the
// compiler regroups the sum and packs with a v_lshl_or_b32; this form packs
// the upper count with a v_lshlrev_b32, which issues at twice the rate.
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}
__device__ __forceinline__ uint32_t popc_acc(uint64_t a, uint32_t acc) {
    return bcnt_acc((uint32_t)a, bcnt_acc((uint32_t)(a >> 32), acc));
}

// One recurrence step of a line, NP packed u16 pairs (2 NP disparities) per lane:
//   L(d) = C(d) + min(L'(d), B-(d), B+(d), m' + P2) - m'
//   B-(d) = d == 0 ? 127 : L'(d-1) + P1,  B+(d) = d == D-1 ? 127 : L'(d+1) + P1
template <int NP>
struct Line {
    uint32_t L[NP];
    uint32_t m;   // the line minimum in both u16 halves
};

// both halves = the smaller half (v_pk_min_u16 with swapped operand halves);
// a u32 minimum of such values is their u16 minimum
__device__ __forceinline__ uint32_t pk_min_halves(uint32_t v) {
    return as_u(__builtin_elementwise_min(as_v(v), as_v(v).yx));
}

template <int LPL, int NP>
__device__ __forceinline__ void line_init(Line<NP>& s, const uint32_t (&c)[NP]) {
    uint32_t t = c[0];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        s.L[k] = c[k];
        t = pk_min(t, c[k]);
    }
    s.m = line_min<LPL>(pk_min_halves(t));
}

template <int LPL, int NP>
__device__ __forceinline__ void line_step(Line<NP>& s, const uint32_t (&c)[NP], uint32_t P1v,
                                          uint32_t P2v, bool first, bool last) {
    const uint32_t mv = s.m;
    const uint32_t mp2 = pk_add(mv, P2v);
    uint32_t R[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) R[k] = pk_add(s.L[k], P1v);
    const uint32_t prev_last = from_prev<LPL>(R[NP - 1], 127u << 16, first);
    const uint32_t next_first = from_next<LPL>(R[0], 127u, last);
    uint32_t E[NP + 1];
    E[0] = __builtin_amdgcn_alignbit(R[0], prev_last, 16);
#pragma unroll
    for (int k = 1; k < NP; ++k) E[k] = __builtin_amdgcn_alignbit(R[k], R[k - 1], 16);
    E[NP] = __builtin_amdgcn_alignbit(next_first, R[NP - 1], 16);
    uint32_t T[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const uint32_t b = pk_min(E[k], E[k + 1]);
        const uint32_t v = pk_min(pk_min(s.L[k], b), mp2);
        s.L[k] = pk_sub(pk_add(c[k], v), mv);
        T[k] = s.L[k];
    }
    // tree minimum over the pairs
#pragma unroll
    for (int w = 1; w < NP; w *= 2)
#pragma unroll
        for (int k = 0; k + w < NP; k += 2 * w) T[k] = pk_min(T[k], T[k + w]);
    s.m = line_min<LPL>(pk_min_halves(T[0]));
}

// the lane's 2 NP bytes (uint8 truncation) in memory order; `sel` reverses
// the bytes of each word, `revmask` (all ones or zero, wave-uniform) the
// order of the words.  The mask select keeps the word order a bit operation:
// a conditional on `rev` is turned by the compiler into a per-step indexed
// select chain.
template <int NP>
__device__ __forceinline__ void pack_lane(const Line<NP>& s, uint32_t sel, uint32_t revmask,
                                          uint32_t (&o)[NP / 2]) {
    uint32_t w[NP / 2];
#pragma unroll
    for (int k = 0; k < NP / 2; ++k) w[k] = __builtin_amdgcn_perm(s.L[2 * k + 1], s.L[2 * k], sel);
#pragma unroll
    for (int k = 0; k < NP / 2; ++k) o[k] = (w[NP / 2 - 1 - k] & revmask) | (w[k] & ~revmask);
}

constexpr uint32_t kSelFwd = 0x06040200u;   // bytes d, d+1, d+2, d+3
constexpr uint32_t kSelRev = 0x00020406u;   // the same bytes, reversed

__device__ __forceinline__ uint64_t g64(const uint64_t* p) { return gld(p); }

__device__ __forceinline__ int clampcol(int p, int W) { return (p >= 0 && p < W) ? p : 0; }

// value of lane lane0 + SRC (SRC < LPL) of the line, on every lane of the line
template <int LPL>
__device__ __forceinline__ uint32_t line_bcast(uint32_t v, int src, int lane0) {
    if constexpr (LPL == 4) {
        switch (src) {   // quad_perm [src, src, src, src]
            case 0: return dpp0<0x00>(v);
            case 1: return dpp0<0x55>(v);
            case 2: return dpp0<0xAA>(v);
            default: return dpp0<0xFF>(v);
        }
    } else {
        return (uint32_t)__builtin_amdgcn_ds_bpermute((lane0 + src) << 2, (int)v);
    }
}
template <int LPL>
__device__ __forceinline__ uint64_t line_bcast64(uint64_t v, int src, int lane0) {
    const uint32_t lo = line_bcast<LPL>((uint32_t)v, src, lane0);
    const uint32_t hi = line_bcast<LPL>((uint32_t)(v >> 32), src, lane0);
    return ((uint64_t)hi << 32) | lo;
}

// packed costs of one lane: pair k = (partner 2k, partner 2k + 1)
template <int NP, class F>
__device__ __forceinline__ void lane_costs(uint64_t own, F&& partner, uint32_t (&c)[NP]) {
#pragma unroll
    for (int k = 0; k < NP; ++k)
        c[k] = popc_acc(own ^ partner(2 * k), popc64(own ^ partner(2 * k + 1)) << 16);
}

// ------------------------------------------------------------ horizontal ---
// Left->right or right->left over the wave's rows.  With t' the lane-local
// order (t' = t, or D-1-t when reversed), the partner column is
//   p(x, t') = a + x + st * t',   and   p(x + dx, t') = p(x, t' - 1),
// so every step the window moves up by one: lane q takes lane q-1's element
// t' = DPL q - 1 and the line's lane 0 loads p(x, 0) = a + x.
template <int D>
__device__ __forceinline__ void wide_horizontal(const SgmJob& J, int path, int group, int lane) {
    constexpr int DPL = wide_dpl(kWideH, D), NP = DPL / 2, LPL = D / DPL, LPW = 64 / LPL;
    const int q = lane % LPL, sub = lane / LPL;
    const bool first = q == 0, last = q == LPL - 1;
    const int W = J.W, H = J.H, roi_x = J.roi_x, min_d = J.min_d;
    int y = group * LPW + sub;
    const bool active = y < H;
    if (!active) y = H - 1;
    const bool fwd = path == kLR;
    const int n = W - roi_x;
    const int x0 = fwd ? roi_x : W - 1, dx = fwd ? 1 : -1;
    const bool right = J.is_right != 0;
    const bool rev = right != (path == kRL);
    const int a = right ? (rev ? min_d + D - 1 : min_d) : (rev ? -min_d - (D - 1) : -min_d);
    const int st = (right != rev) ? 1 : -1;
    const uint64_t* cb = J.census + (int64_t)y * W;
    const uint64_t* cm = J.census_other + (int64_t)y * W;
    const uint32_t P1v = (uint32_t)J.p1 * 0x10001u, P2v = (uint32_t)J.p2 * 0x10001u;
    const uint32_t sel = rev ? kSelRev : kSelFwd;
    const uint32_t revmask = rev ? ~0u : 0u;
    const int chunk = rev ? LPL - 1 - q : q;
    uint8_t* const out_base = (path == kLR) ? J.path[kLR] : J.path[kRL];
    uint8_t* const dst0 = out_base + ((int64_t)y * W + x0) * D + chunk * DPL;
    const int64_t dstep = (int64_t)dx * D;

    // window: slot j holds element t' = DPL q + j at the start of a block
    uint64_t win[DPL];
#pragma unroll
    for (int j = 0; j < DPL; ++j) win[j] = g64(cm + clampcol(a + x0 + st * (DPL * q + j), W));
    Line<NP> s;
    {
        uint32_t c[NP];
        lane_costs<NP>(g64(cb + x0), [&](int j) { return win[j]; }, c);
        line_init<LPL>(s, c);
        uint32_t o[NP / 2];
        pack_lane(s, sel, revmask, o);
        if (active) store_words<NP / 2>(dst0, o);
    }

    // Census words of a block of DPL steps, fetched one block ahead: lane q
    // of a line loads the own and entering-partner words of steps
    // q + LPL k (k < E), and each step takes them from the line's lane
    // (step mod LPL) by a lane broadcast.  Few, early loads keep the census
    // traffic (many scattered 8-byte reads) off the step's critical path.
    constexpr int E = DPL / LPL;
    static_assert(E * LPL == DPL, "a block's steps spread evenly over the line's lanes");
    uint64_t c_own[E], c_in[E], n_own[E], n_in[E];   // this block, next block
    auto xat = [&](int i) { return x0 + dx * min(i, n - 1); };
    auto load_block = [&](int i0, uint64_t (&o)[E], uint64_t (&in)[E]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int x = xat(i0 + q + LPL * k);
            o[k] = g64(cb + x);
            in[k] = g64(cm + clampcol(a + x, W));
        }
    };
    const int lane0 = lane - q;   // first lane of the line
    load_block(1, c_own, c_in);
    // blocks of DPL steps: the window slots rotate once per block
    for (int i0 = 1; i0 < n; i0 += DPL) {
        load_block(i0 + DPL, n_own, n_in);
#pragma unroll
        for (int sb = 0; sb < DPL; ++sb) {
            const int i = i0 + sb;
            const uint64_t own = line_bcast64<LPL>(c_own[sb / LPL], sb % LPL, lane0);
            const uint64_t in = line_bcast64<LPL>(c_in[sb / LPL], sb % LPL, lane0);
            // the slot of the leaving element t' = DPL q + DPL - 1 takes the
            // entering one; element j then sits in slot (e + j) mod DPL
            const int e = (DPL - 1 - sb) & (DPL - 1);
            const uint64_t lv = win[e];
            const uint32_t nlo = from_prev<LPL>((uint32_t)lv, (uint32_t)in, first);
            const uint32_t nhi = from_prev<LPL>((uint32_t)(lv >> 32), (uint32_t)(in >> 32), first);
            win[e] = ((uint64_t)nhi << 32) | nlo;
            uint32_t c[NP];
            lane_costs<NP>(own, [&](int j) { return win[(e + j) & (DPL - 1)]; }, c);
            line_step<LPL>(s, c, P1v, P2v, first, last);
            uint32_t o[NP / 2];
            pack_lane(s, sel, revmask, o);
            if (active && i < n) store_words<NP / 2>(dst0 + (int64_t)i * dstep, o);
        }
#pragma unroll
        for (int k = 0; k < E; ++k) {
            c_own[k] = n_own[k];
            c_in[k] = n_in[k];
        }
    }
}

// -------------------------------------------------------------- vertical ---
// Top->bottom (stored) or bottom->top fused with the WTA, over the wave's
// columns.  Per step the wave stages NV = LPW + D + LPW - 1 census words in
// LDS: [0, LPW) its own columns, then the partner columns pbase + k.
template <int D, int DPL>
struct VertGeom {
    static constexpr int LPL = D / DPL, LPW = 64 / LPL;
    static constexpr int NV = D + 2 * LPW - 1;
    static constexpr int NR = (NV + 63) / 64;   // words loaded per lane per step
};

template <int DMAX, int KIND>
struct VertLds {
    static constexpr int NV = (DMAX >= 256) ? VertGeom<256, wide_dpl(KIND, 256)>::NV
                            : (DMAX >= 128) ? VertGeom<128, wide_dpl(KIND, 128)>::NV
                                            : VertGeom<64, wide_dpl(KIND, 64)>::NV;
};

template <int D, bool FUSE>
__device__ __forceinline__ void wide_vertical(const SgmJob& J, int group, int lane, uint64_t* lds,
                                              uint16_t* sums /* FUSE: [lines][D] */) {
    constexpr int KIND = FUSE ? kWideF : kWideV;
    constexpr int DPL = wide_dpl(KIND, D), NP = DPL / 2, NW = NP / 2;
    using G = VertGeom<D, DPL>;
    constexpr int LPL = G::LPL, LPW = G::LPW, NV = G::NV, NR = G::NR;
    const int q = lane % LPL, sub = lane / LPL;
    const bool first = q == 0, last = q == LPL - 1;
    const int W = J.W, H = J.H, roi_x = J.roi_x, min_d = J.min_d;
    const int nlines = W - roi_x;
    const int line0 = group * LPW;
    int line = line0 + sub;
    const bool active = line < nlines;
    if (!active) line = nlines - 1;
    const int c0 = roi_x + line0, x = roi_x + line;
    const int y0 = FUSE ? H - 1 : 0, dy = FUSE ? -1 : 1;
    const bool right = J.is_right != 0;
    const int pbase = right ? c0 + min_d : c0 - min_d - (D - 1);
    // this lane's LDS indices: own word, then partner t = DPL q + j at
    // kp + j (right view) or kp - j (left view)
    const int kown = x - c0;
    const int kp = LPW + (right ? (x - c0) + DPL * q : (x - c0) + (D - 1) - DPL * q);
    const uint32_t P1v = (uint32_t)J.p1 * 0x10001u, P2v = (uint32_t)J.p2 * 0x10001u;
    // volume offsets: a wave-uniform row part and the lane's 32-bit part
    const int64_t ubase = ((int64_t)y0 * W + c0) * D;
    const uint32_t loff = (uint32_t)((x - c0) * D + q * DPL);
    const int64_t off0 = ubase + loff;
    const int64_t ostep = dy * (int64_t)W * D;
    uint8_t* const out_base = FUSE ? J.path[kDU] : J.path[kUD];

    // census words staged for step i: lane loads v = lane + 64 r
    auto load_step = [&](int i, uint64_t (&g)[NR]) __attribute__((always_inline)) {
        const int y = y0 + dy * min(i, H - 1);
        const uint64_t* cb = J.census + (int64_t)y * W;
        const uint64_t* cm = J.census_other + (int64_t)y * W;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int v = min(lane + 64 * r, NV - 1);
            g[r] = (v < LPW) ? g64(cb + min(c0 + v, W - 1)) : g64(cm + clampcol(pbase + v - LPW, W));
        }
    };
    auto stage = [&](int slot, const uint64_t (&g)[NR]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const int v = lane + 64 * r;
            if (v < NV) lds[slot * NV + v] = g[r];
        }
    };
    auto costs = [&](int slot, uint32_t (&c)[NP]) __attribute__((always_inline)) {
        const uint64_t* b = lds + slot * NV;
        lane_costs<NP>(b[kown], [&](int j) { return right ? b[kp + j] : b[kp - j]; }, c);
    };

    // WTA (FUSE): lane q < 4 of a line keeps the integer result of step
    // i0 + q of the 4-step block i0 .. i0 + 3 (i0 = 1 + 4b); the block ends
    // with those lanes finishing their pixels with the float rules.
    const float uthr = 1.0f - J.uniqueness;
    uint32_t kb = 0u, kboth = 0u;
    uint16_t* const sline = sums + sub * D;   // this line's sums (FUSE)
    const uint8_t* srcs[3] = {J.path[kUD], J.path[kLR], J.path[kRL]};
    auto wta_flush = [&](int i0) __attribute__((always_inline)) {
        const int i = i0 + q;
        const int y = y0 - i;
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
            const float fa = c1 - c2f, fb = c3 - c2f;
            const float uniq = (fa < fb) ? fa : fb;
            if (!(uniq <= c2f * uthr)) {
                const float t = c1 + c3 - 2.0f * c2f;
                const float den = (1.0f < t) ? t : 1.0f;
                const float bias = (float)((double)(c1 - c3 + den) / (double)(2.0f * den));
                outv = (float)(ds + min_d) + bias;
            }
        }
        if (active && q < 4 && i < H) gst(J.disp + (int64_t)y * W + x, outv);
    };
    auto wta_step = [&](const Line<NP>& s, const uint32_t (&o3)[3][NW], int i) __attribute__((always_inline)) {
        uint32_t S[NP];
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const int w = k >> 1;
            const uint32_t a0 = (k & 1) ? bytes23(o3[0][w]) : bytes01(o3[0][w]);
            const uint32_t a1 = (k & 1) ? bytes23(o3[1][w]) : bytes01(o3[1][w]);
            const uint32_t a2 = (k & 1) ? bytes23(o3[2][w]) : bytes01(o3[2][w]);
            S[k] = a0 + a1 + a2 + (s.L[k] & 0x00FF00FFu);   // each half <= 1020
        }
        // argmin with the smallest d: inside the lane on u16 keys S * DPL + j
        // (S <= 4 * 255 < 2^10 and j < DPL <= 64: the keys fit in 16 bits, so
        // one 32-bit v_lshl_or_b32 per pair shifts both halves without a carry
        // between them), then across the line's lanes on (S << 16) | d
        static_assert(DPL <= 64 && (DPL & (DPL - 1)) == 0, "u16 argmin keys S * DPL + j");
        constexpr int kLog = __builtin_ctz(DPL);
        uint32_t key = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            const uint32_t jk = (uint32_t)(2 * k) | ((uint32_t)(2 * k + 1) << 16);
            key = pk_min(key, (S[k] << kLog) | jk);
        }
        key = pk_min_halves(key) & 0xFFFFu;
        uint32_t best = ((key / DPL) << 16) | (uint32_t)(DPL * q + key % DPL);
        best = line_min<LPL>(best);
        const int ds = (int)(best & 0xFFFFu);
        // S(ds - 1) and S(ds + 1): the lanes of a line hold consecutive
        // disparities, so their packed sums laid side by side in LDS are the
        // line's u16 S[d]; every lane then reads the two it needs.  Indices
        // are clamped: S(-1) and S(D) are never used (ds = 0 is invalid, and
        // ds = D - 1 takes c3 = D - 1).
        uint32_t* const mine = reinterpret_cast<uint32_t*>(sline + DPL * q);
#pragma unroll
        for (int k = 0; k < NP; ++k) mine[k] = S[k];
        __builtin_amdgcn_wave_barrier();
        const uint32_t c1 = sline[max(ds - 1, 0)], c3 = sline[min(ds + 1, D - 1)];
        __builtin_amdgcn_wave_barrier();
        if (q == ((i - 1) & 3)) { kb = best; kboth = (c1 << 16) | c3; }
    };

    // rings (PF steps ahead) of census words and, with FUSE, of the other
    // three directions' bytes; the loop body covers BLK steps (with FUSE one
    // 4-step WTA block)
    constexpr int PF = FUSE ? FSGM_PF_F : FSGM_PF_V;
    constexpr int BLK = (PF > 4) ? PF : 4;
    static_assert(BLK % PF == 0, "the loop body must hold whole ring turns");
    static_assert(!FUSE || BLK == 4, "the WTA finishes its pixels in 4-step blocks");
    uint64_t gr[PF][NR];
    uint32_t pr[FUSE ? PF : 1][3][NW];
    {
        uint64_t g0[NR];
        load_step(0, g0);
        stage(0, g0);
    }
    // the prologue issues the rings in the loop's own order (the wait counts
    // at the loop head merge both paths)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int p = 0; p < PF; ++p) {
        load_step(1 + p, gr[p]);
        if constexpr (FUSE) {
#pragma unroll
            for (int u = 0; u < 3; ++u)
                load_words<NW>(srcs[u] + off0 + (int64_t)min(1 + p, H - 1) * ostep, pr[p][u]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_wave_barrier();
    Line<NP> s;
    {
        uint32_t c[NP];
        costs(0, c);
        line_init<LPL>(s, c);
        if constexpr (!FUSE) {
            uint32_t o[NW];
            pack_lane(s, kSelFwd, 0u, o);
            if (active) store_words<NW>(out_base + off0, o);
        } else {
            // row H - 1 is a border row of the WTA
            if (active && first) gst(J.disp + (int64_t)y0 * W + x, kInvalidDisp);
        }
    }
    // row offsets of step i (stores) and of the refill step min(i + PF, H - 1)
    // (loads), advanced by scalar adds: no 64-bit multiply per access
    int64_t row_i = ubase + ostep, row_f = ubase + (int64_t)(1 + PF) * ostep;
    const int64_t row_last = ubase + (int64_t)(H - 1) * ostep;
    for (int i0 = 1; i0 < H; i0 += BLK) {
#pragma unroll
        for (int b = 0; b < BLK; ++b) {
            const int i = i0 + b, p = b % PF;
            // stage step i (slot i & 1); the other slot still holds step i - 1
            stage(i & 1, gr[p]);
            load_step(i + PF, gr[p]);
            __builtin_amdgcn_wave_barrier();
            uint32_t c[NP];
            costs(i & 1, c);
            line_step<LPL>(s, c, P1v, P2v, first, last);
            if constexpr (!FUSE) {
                uint32_t o[NW];
                pack_lane(s, kSelFwd, 0u, o);
                int64_t r = row_i;
                asm volatile("" : "+s"(r));   // keep it a scalar of its own
                if (active && i < H) store_words<NW>(out_base + r + loff, o);
            } else {
                // consume the slot, then refill it for step i + PF
                wta_step(s, pr[p], i);
                int64_t r = (i + PF <= H - 1) ? row_f : row_last;
                asm volatile("" : "+s"(r));
#pragma unroll
                for (int u = 0; u < 3; ++u) load_words<NW>(srcs[u] + r + loff, pr[p][u]);
            }
            row_i += ostep;
            row_f += ostep;
        }
        if constexpr (FUSE) wta_flush(i0);
    }
}

// Kernels.  The horizontal (left->right, right->left) and vertical
// (top->bottom, or bottom->top fused with the WTA) scanline code run as
// separate kernels: each fits 4+ waves per SIMD on its own, together they
// would not.  Grid form: item `local` of a job is its scanline group
// (horizontal: left->right groups, then right->left groups).
template <int DMAX>
__global__ __launch_bounds__(64) void wide_h_kernel(const SgmJob* __restrict__ jobs,
                                                    const int* __restrict__ offs, int njobs) {
    const int item = blockIdx.x;
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const int local = item - offs[j];
    const int lw = wide_lines(kWideH, J.D);
    const int nh = (J.H + lw - 1) / lw;
    const int path = local < nh ? kLR : kRL, group = local < nh ? local : local - nh;
    switch (J.D) {
        case 64: wide_horizontal<64>(J, path, group, threadIdx.x); break;
        case 128: if constexpr (DMAX >= 128) wide_horizontal<128>(J, path, group, threadIdx.x); break;
        case 256: if constexpr (DMAX >= 256) wide_horizontal<256>(J, path, group, threadIdx.x); break;
        default: break;
    }
}

// sums of the fused WTA: 64 * DPL u16 per wave (lines x D), largest at DMAX
template <int DMAX>
constexpr int kSumsLen = 64 * wide_dpl(kWideF, DMAX);

template <int DMAX, bool FUSE>
__global__ __launch_bounds__(64) void wide_v_kernel(const SgmJob* __restrict__ jobs,
                                                    const int* __restrict__ offs, int njobs) {
    __shared__ uint64_t lds[2 * VertLds<DMAX, FUSE ? kWideF : kWideV>::NV];
    __shared__ __align__(16) uint16_t sums[FUSE ? kSumsLen<DMAX> : 1];
    const int item = xcd_item(blockIdx.x, gridDim.x);
    const int j = find_job(offs, njobs, item);
    const SgmJob& J = jobs[j];
    const int group = item - offs[j];
    switch (J.D) {
        case 64: wide_vertical<64, FUSE>(J, group, threadIdx.x, lds, sums); break;
        case 128: if constexpr (DMAX >= 128) wide_vertical<128, FUSE>(J, group, threadIdx.x, lds, sums); break;
        case 256: if constexpr (DMAX >= 256) wide_vertical<256, FUSE>(J, group, threadIdx.x, lds, sums); break;
        default: break;
    }
}

// Persistent forms: packed items (job << 16 | direction << 14 | group)
// pulled from a device counter by a fixed grid of resident waves; every wave
// leaves through the one test on the counter.
__device__ __forceinline__ bool next_item(const uint32_t* items, int n, uint32_t* counter, uint32_t* it) {
    uint32_t idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(counter, 1u);
    idx = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(idx, 0));
    if (idx >= (uint32_t)n) return false;
    *it = items[idx];
    return true;
}

template <int DMAX>
__global__ __launch_bounds__(64) void wide_h_persistent(const SgmJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ items, int n,
                                                        uint32_t* __restrict__ counter) {
    uint32_t it;
    while (next_item(items, n, counter, &it)) {
        const SgmJob& J = jobs[it >> 16];
        const int path = (int)((it >> 14) & 3u), group = (int)(it & 0x3FFFu);
        switch (J.D) {
            case 64: wide_horizontal<64>(J, path, group, threadIdx.x); break;
            case 128: if constexpr (DMAX >= 128) wide_horizontal<128>(J, path, group, threadIdx.x); break;
            case 256: if constexpr (DMAX >= 256) wide_horizontal<256>(J, path, group, threadIdx.x); break;
            default: break;
        }
    }
}

template <int DMAX, bool FUSE>
__global__ __launch_bounds__(64) void wide_v_persistent(const SgmJob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ items, int n,
                                                        uint32_t* __restrict__ counter) {
    __shared__ uint64_t lds[2 * VertLds<DMAX, FUSE ? kWideF : kWideV>::NV];
    __shared__ __align__(16) uint16_t sums[FUSE ? kSumsLen<DMAX> : 1];
    uint32_t it;
    while (next_item(items, n, counter, &it)) {
        const SgmJob& J = jobs[it >> 16];
        const int group = (int)(it & 0x3FFFu);
        switch (J.D) {
            case 64: wide_vertical<64, FUSE>(J, group, threadIdx.x, lds, sums); break;
            case 128: if constexpr (DMAX >= 128) wide_vertical<128, FUSE>(J, group, threadIdx.x, lds, sums); break;
            case 256: if constexpr (DMAX >= 256) wide_vertical<256, FUSE>(J, group, threadIdx.x, lds, sums); break;
            default: break;
        }
    }
}

hipError_t launch_h(const StagePlan& p, hipStream_t s) {
    if (p.n_wh <= 0) return hipSuccess;
    const dim3 g(p.n_wh), b(64);
    if (p.dmax <= 64) hipLaunchKernelGGL(wide_h_kernel<64>, g, b, 0, s, p.jobs, p.offs_wh, p.njobs);
    else if (p.dmax <= 128) hipLaunchKernelGGL(wide_h_kernel<128>, g, b, 0, s, p.jobs, p.offs_wh, p.njobs);
    else hipLaunchKernelGGL(wide_h_kernel<256>, g, b, 0, s, p.jobs, p.offs_wh, p.njobs);
    return hipGetLastError();
}

template <bool FUSE>
hipError_t launch_v(const StagePlan& p, hipStream_t s) {
    const int n = FUSE ? p.n_wf : p.n_wv;
    const int* offs = FUSE ? p.offs_wf : p.offs_wv;
    if (n <= 0) return hipSuccess;
    const dim3 g(n), b(64);
    if (p.dmax <= 64)
        hipLaunchKernelGGL((wide_v_kernel<64, FUSE>), g, b, 0, s, p.jobs, offs, p.njobs);
    else if (p.dmax <= 128)
        hipLaunchKernelGGL((wide_v_kernel<128, FUSE>), g, b, 0, s, p.jobs, offs, p.njobs);
    else
        hipLaunchKernelGGL((wide_v_kernel<256, FUSE>), g, b, 0, s, p.jobs, offs, p.njobs);
    return hipGetLastError();
}

hipError_t launch_h_persistent(const StagePlan& p, const uint32_t* items, int n, uint32_t* counter,
                               int grid, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 g(grid), b(64);
    if (p.dmax <= 64) hipLaunchKernelGGL(wide_h_persistent<64>, g, b, 0, s, p.jobs, items, n, counter);
    else if (p.dmax <= 128) hipLaunchKernelGGL(wide_h_persistent<128>, g, b, 0, s, p.jobs, items, n, counter);
    else hipLaunchKernelGGL(wide_h_persistent<256>, g, b, 0, s, p.jobs, items, n, counter);
    return hipGetLastError();
}

template <bool FUSE>
hipError_t launch_v_persistent(const StagePlan& p, const uint32_t* items, int n, uint32_t* counter,
                               int grid, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 g(grid), b(64);
    if (p.dmax <= 64)
        hipLaunchKernelGGL((wide_v_persistent<64, FUSE>), g, b, 0, s, p.jobs, items, n, counter);
    else if (p.dmax <= 128)
        hipLaunchKernelGGL((wide_v_persistent<128, FUSE>), g, b, 0, s, p.jobs, items, n, counter);
    else
        hipLaunchKernelGGL((wide_v_persistent<256, FUSE>), g, b, 0, s, p.jobs, items, n, counter);
    return hipGetLastError();
}

}  // namespace

// left->right + right->left, then top->bottom (one stream, in order)
hipError_t launch_wide3(const StagePlan& p, hipStream_t s) {
    const hipError_t e = launch_h(p, s);
    return e != hipSuccess ? e : launch_v<false>(p, s);
}

hipError_t launch_wide_wta(const StagePlan& p, hipStream_t s) { return launch_v<true>(p, s); }

// schedule: items3 = [horizontal items (n3h)][top->bottom items]; counters
// [0] horizontal, [1] bottom->top, [2] top->bottom
hipError_t launch_wide3_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s) {
    hipError_t e = launch_reset_counters(sc.counters, 3, s);
    if (e != hipSuccess) return e;
    if ((e = launch_h_persistent(p, sc.items3, sc.n3h, sc.counters, sc.grid, s)) != hipSuccess) return e;
    return launch_v_persistent<false>(p, sc.items3 + sc.n3h, sc.n3 - sc.n3h, sc.counters + 2, sc.grid, s);
}

hipError_t launch_wide_wta_sched(const StagePlan& p, const AggSchedule& sc, hipStream_t s) {
    return launch_v_persistent<true>(p, sc.items_dw, sc.ndw, sc.counters + 1, sc.grid, s);
}

// resident waves per CU: the smallest over the three persistent kernels
hipError_t wide_occupancy(const StagePlan& p, int* per_cu) {
    int a = 0, b = 0, c = 0;
    hipError_t e;
    if (p.dmax <= 64) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, wide_h_persistent<64>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, wide_v_persistent<64, false>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, wide_v_persistent<64, true>, 64, 0);
    } else if (p.dmax <= 128) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, wide_h_persistent<128>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, wide_v_persistent<128, false>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, wide_v_persistent<128, true>, 64, 0);
    } else {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, wide_h_persistent<256>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, wide_v_persistent<256, false>, 64, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, wide_v_persistent<256, true>, 64, 0);
    }
    if (e == hipSuccess) *per_cu = std::max(1, std::min(a, std::min(b, c)));
    return e;
}

}  // namespace fsgm
