// Differentiable FP32 SGM for CDNA4 (gfx950): forward, soft-argmin and a
// hand-written backward.
//
// Definition (per image b of a batch; cost C[B][H][W][D], per-pixel penalty
// maps P1[B][H][W], P2[B][H][W], temperature T):
//   per direction r (0 top->bottom, 1 left->right, 2 bottom->top,
//   3 right->left) over the whole image
//     L(p0, d) = C(p0, d)
//     L(p, d)  = C(p, d) + min(L'(d), min(L'(d-1), L'(d+1)) + P1(p), m' + P2(p)) - m'
//     L'(-1) = L'(D) = +inf,  m' = min_d L'(d)
//   S = ((L_1 + L_3) + L_0) + L_2
//   disp(p) = sum_d d * softmax_d(-S(p, d) / T)
// Subgradients: the first min candidate in the order [L'(d), L'(d-1)+P1,
// L'(d+1)+P1, m'+P2] wins; m' = L'(k*) with k* the smallest argmin.
//
// Memory: S is built in place and no per-direction volume is kept.  The
// forward's first launch runs, per group of 4 rows, left->right (S = L_1)
// and then right->left (S += L_3) in one wave, so the second pass adds onto
// the first with no other writer; beside them the top->bottom groups write
// L_0 into the one workspace volume.  The second launch runs bottom->top,
// forms S = (S + L_0) + L_2 and the soft-argmin, and keeps the minimum of S
// and the softmax normaliser per pixel.  Every direction also saves, per
// step, the winner code of each disparity (2 bits) and k* of the previous
// step's L (one byte per pixel): 1 byte per cell over the four directions,
// a quarter of a volume, which is all the reverse recurrence needs of the
// forward.  The backward walks each scanline from its last step to its
// first, once:
//   G(s-1, d) = gS(s-1, d) + G(s, d)[w(s,d)=0] + G(s, d+1)[w(s,d+1)=1]
//             + G(s, d-1)[w(s,d-1)=2] + [d = k*(s-1)] (sum_e G(s,e)[w(s,e)=3] - sum_e G(s,e))
// with gS = dLoss/dS formed from S, the kept minimum and normaliser, and
// emits dC_r(s, d) = G(s, d), dP1_r(s) = sum_d G(s,d)[w in {1,2}],
// dP2_r(s) = sum_d G(s,d)[w = 3].  Its launches mirror the forward's: dC,
// dP1 and dP2 are built in place by the horizontal pairs, the top->bottom
// terms go to the workspace, and bottom->top adds both in the same order as S.
//
// Layout per wave as in the integer aggregation: four 16-lane DPP rows, one
// scanline per row, D/16 disparities per lane.
#include <algorithm>
#include <cmath>

#include "dsgm_internal.h"
#include "sgm_device.h"

#pragma clang fp contract(off)

namespace fsgm {
namespace {

template <int CTRL>
__device__ __forceinline__ float dppf(float old, float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
        __builtin_bit_cast(int, old), __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dppi(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xF, 0xF, false);
}
constexpr int kShl1 = 0x101, kShr1 = 0x111;

__device__ __forceinline__ float row_minf(float v) {
    v = fminf(v, dppf<0x128>(0.f, v));
    v = fminf(v, dppf<0x124>(0.f, v));
    v = fminf(v, dppf<0x122>(0.f, v));
    v = fminf(v, dppf<0x121>(0.f, v));
    return v;
}
__device__ __forceinline__ float row_sumf(float v) {
    v = v + dppf<0x128>(0.f, v);
    v = v + dppf<0x124>(0.f, v);
    v = v + dppf<0x122>(0.f, v);
    v = v + dppf<0x121>(0.f, v);
    return v;
}
__device__ __forceinline__ int row_mini(int v) {
    v = min(v, dppi<0x128>(0, v));
    v = min(v, dppi<0x124>(0, v));
    v = min(v, dppi<0x122>(0, v));
    v = min(v, dppi<0x121>(0, v));
    return v;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) f32x4 g_f4;
typedef __attribute__((address_space(1))) float g_f;

template <int N>
__device__ __forceinline__ void loadf(const float* p, float (&v)[N]) {
#pragma unroll
    for (int k = 0; k < N; k += 4) {
        const f32x4 t = *(const g_f4*)(p + k);
        v[k] = t.x; v[k + 1] = t.y; v[k + 2] = t.z; v[k + 3] = t.w;
    }
}
template <int N>
__device__ __forceinline__ void storef(float* p, const float (&v)[N]) {
#pragma unroll
    for (int k = 0; k < N; k += 4) *(g_f4*)(p + k) = f32x4{v[k], v[k + 1], v[k + 2], v[k + 3]};
}
__device__ __forceinline__ float loads(const float* p) { return *(const g_f*)p; }

// A lane's winner codes of one step: DPD = D / 16 disparities, 2 bits each.
template <int D> struct WinWordOf;
template <> struct WinWordOf<64> { using type = uint8_t; };
template <> struct WinWordOf<128> { using type = uint16_t; };
template <> struct WinWordOf<256> { using type = uint32_t; };
template <int D> using WinWord = typename WinWordOf<D>::type;

// Geometry of one scanline of one direction.
struct LineGeo {
    int n;            // steps
    int line;         // row (horizontal) or column (vertical)
    bool active;
    int64_t pix0;     // pixel index of step 0 within the image
    int64_t pstep;    // pixel index increment per step
};

__device__ __forceinline__ LineGeo line_geo(const DsgmArgs& A, int path, int group, int sub) {
    LineGeo g;
    const bool horiz = (path == 1 || path == 3);
    const bool fwd = (path == 0 || path == 1);
    const int H = A.H, W = A.W;
    const int nlines = horiz ? H : W;
    int line = group * 4 + sub;
    g.active = line < nlines;
    if (!g.active) line = nlines - 1;
    g.line = line;
    g.n = horiz ? W : H;
    if (horiz) {
        g.pix0 = (int64_t)line * W + (fwd ? 0 : W - 1);
        g.pstep = fwd ? 1 : -1;
    } else {
        g.pix0 = (int64_t)(fwd ? 0 : H - 1) * W + line;
        g.pstep = fwd ? W : -W;
    }
    return g;
}

// one forward step for a lane's DPD disparities; returns winners packed 2 bits each
template <int DPD>
__device__ __forceinline__ uint32_t fwd_step(float (&L)[DPD], const float (&c)[DPD], float m,
                                             float p1, float p2) {
    const float inf = __builtin_inff();
    const float prev_last = dppf<kShr1>(inf, L[DPD - 1]);
    const float next_first = dppf<kShl1>(inf, L[0]);
    const float c3 = m + p2;
    uint32_t wins = 0;
    float nl[DPD];
#pragma unroll
    for (int k = 0; k < DPD; ++k) {
        const float lm = (k == 0) ? prev_last : L[k - 1];
        const float lp = (k == DPD - 1) ? next_first : L[k + 1];
        float best = L[k];
        uint32_t w = 0;
        const float c1 = lm + p1, c2 = lp + p1;
        if (c1 < best) { best = c1; w = 1; }
        if (c2 < best) { best = c2; w = 2; }
        if (c3 < best) { best = c3; w = 3; }
        nl[k] = (c[k] + best) - m;
        wins |= w << (2 * k);
    }
#pragma unroll
    for (int k = 0; k < DPD; ++k) L[k] = nl[k];
    return wins;
}

template <int DPD>
__device__ __forceinline__ float lane_min(const float (&L)[DPD]) {
    float v = L[0];
#pragma unroll
    for (int k = 1; k < DPD; ++k) v = fminf(v, L[k]);
    return v;
}

// smallest d with L(d) == m over the row
template <int DPD>
__device__ __forceinline__ int row_argmin(const float (&L)[DPD], float m, int l) {
    int idx = 0x7FFFFFFF;
#pragma unroll
    for (int k = DPD - 1; k >= 0; --k)
        if (L[k] == m) idx = l * DPD + k;
    return row_mini(idx);
}

// ------------------------------------------------------------- forward ---
// Unnormalised softmax weight exp(-(s - mn) / T) = 2^((mn - s) * c) with
// c = log2(e) / T: one v_exp_f32 after a subtraction and a product.  The
// forward and the backward use this same form.  The argument is <= 0, and a
// weight below 2^-126 (a cost more than 87 T above the minimum) contributes
// nothing to the sums either way.
__device__ __forceinline__ float soft_weight(float s, float mn, float c) {
    return __builtin_amdgcn_exp2f((mn - s) * c);
}
__device__ __forceinline__ float soft_scale(float invT) { return invT * 1.44269504088896340736f; }

// Soft-argmin of one pixel per 16-lane row: s holds S, every lane of the row
// receives the disparity, the minimum mn and the normaliser z.
template <int DPD>
__device__ __forceinline__ float soft_argmin(const float (&s)[DPD], int l, float invT, float* mn_out,
                                             float* z_out) {
    const float mn = row_minf(lane_min(s));
    const float c = soft_scale(invT);
    float z = 0.f, acc = 0.f;
#pragma unroll
    for (int k = 0; k < DPD; ++k) {
        const float e = soft_weight(s[k], mn, c);
        z += e;
        acc += (float)(l * DPD + k) * e;
    }
    z = row_sumf(z);
    acc = row_sumf(acc);
    *mn_out = mn;
    *z_out = z;
    return acc / z;
}

// What a forward pass does with each step's L.
enum FwdMode {
    kFwdStoreS = 0,   // S = L          (left->right, first of the horizontal pair)
    kFwdAddS = 1,     // S = S + L      (right->left, second of the pair)
    kFwdStoreY = 2,   // Y = L          (top->bottom, into the workspace volume)
    kFwdLast = 3,     // S = (S + Y) + L, then the soft-argmin (bottom->top)
};

// Loads run ahead of the arithmetic: a scanline step needs its inputs from
// HBM, and a lone wave (the long horizontal pairs run about one per SIMD)
// would otherwise wait the full memory latency at every step.  The inputs of
// step s + kAheadF are issued right after step s has used its slot, into
// register slots indexed by s % kAheadF (compile-time within the unrolled
// block).  Steps past the end load the last pixel again and are never used.
constexpr int kAheadF = 4;

// One direction over a scanline group (forward), saving each step's winner
// codes and the previous step's argmin for the backward.
template <int D, int MODE>
__device__ __forceinline__ void fwd_lines(const DsgmArgs& A, int b, int path, int group, int lane) {
    constexpr int DPD = D / 16;
    constexpr bool kPriorS = MODE == kFwdAddS || MODE == kFwdLast;
    constexpr bool kPriorY = MODE == kFwdLast;
    const int sub = lane >> 4, l = lane & 15;
    const LineGeo g = line_geo(A, path, group, sub);
    const int64_t img = (int64_t)A.H * A.W;
    const float* cost = A.cost + (int64_t)b * img * D + l * DPD;
    const float* p1 = A.p1 + (int64_t)b * img;
    const float* p2 = A.p2 + (int64_t)b * img;
    float* S = A.S + (int64_t)b * img * D + l * DPD;
    float* Y = A.vol + (int64_t)b * img * D + l * DPD;
    float* disp = A.disp + (int64_t)b * img;
    float* stats = A.stats + (int64_t)b * img * 2;
    const int64_t npix = (int64_t)A.B * img;
    // saved for the backward: each step's winner codes (this lane's DPD
    // disparities) and k* of the previous step's L
    WinWord<D>* wout = reinterpret_cast<WinWord<D>*>(A.wins) + ((int64_t)path * npix + (int64_t)b * img) * 16 + l;
    uint8_t* kout = A.kstar + (int64_t)path * npix + (int64_t)b * img;
    const float invT = A.inv_temperature;
    const int n = g.n;

    // inputs of the steps in flight: cost, penalties, and what the pixel's
    // earlier passes left (S, and Y for the last pass)
    float qc[kAheadF][DPD], qp1[kAheadF], qp2[kAheadF];
    float qs[kAheadF][DPD], qy[kAheadF][DPD];
    auto issue = [&](int slot, int s) __attribute__((always_inline)) {
        const int64_t q = g.pix0 + (int64_t)min(s, n - 1) * g.pstep;
        loadf<DPD>(cost + q * D, qc[slot]);
        qp1[slot] = loads(p1 + q);
        qp2[slot] = loads(p2 + q);
        if constexpr (kPriorS) loadf<DPD>(S + q * D, qs[slot]);
        if constexpr (kPriorY) loadf<DPD>(Y + q * D, qy[slot]);
    };
    auto emit = [&](int64_t q, const float (&L)[DPD], int slot) __attribute__((always_inline)) {
        if constexpr (MODE == kFwdLast) {
            float sv[DPD];
#pragma unroll
            for (int k = 0; k < DPD; ++k) sv[k] = (qs[slot][k] + qy[slot][k]) + L[k];
            float mn, z;
            const float d = soft_argmin<DPD>(sv, l, invT, &mn, &z);
            if (g.active) {
                storef<DPD>(S + q * D, sv);
                if (l == 0) {
                    disp[q] = d;
                    stats[2 * q] = mn;
                    stats[2 * q + 1] = z;
                }
            }
        } else if constexpr (MODE == kFwdAddS) {
            float sv[DPD];
#pragma unroll
            for (int k = 0; k < DPD; ++k) sv[k] = qs[slot][k] + L[k];
            if (g.active) storef<DPD>(S + q * D, sv);
        } else {
            if (g.active) storef<DPD>((MODE == kFwdStoreS ? S : Y) + q * D, L);
        }
    };

#pragma unroll
    for (int u = 0; u < kAheadF; ++u) issue(u, u);
    float L[DPD];
    float m = 0.f;
    for (int s0 = 0; s0 < n; s0 += kAheadF) {
#pragma unroll
        for (int u = 0; u < kAheadF; ++u) {
            const int s = s0 + u;
            if (s < n) {
                const int64_t pix = g.pix0 + (int64_t)s * g.pstep;
                if (s == 0) {
#pragma unroll
                    for (int k = 0; k < DPD; ++k) L[k] = qc[u][k];
                } else {
                    const int ks = row_argmin<DPD>(L, m, l);
                    const uint32_t w = fwd_step<DPD>(L, qc[u], m, qp1[u], qp2[u]);
                    if (g.active) {
                        wout[pix * 16] = (WinWord<D>)w;
                        if (l == 0) kout[pix] = (uint8_t)ks;
                    }
                }
                m = row_minf(lane_min(L));
                emit(pix, L, u);
            }
            issue(u, s + kAheadF);
        }
    }
}

// Items of the first launch (forward or backward): the horizontal pairs of
// every image (4 rows each, left->right then right->left in one wave, so the
// second accumulates onto the first without another writer), then the
// top->bottom groups (4 columns each).  The long pairs are dispatched first.
// The second launch is the bottom->top groups.  Column groups are spread so
// that each XCD walks a contiguous run of them: the per-pixel maps (penalties,
// statistics, disparity, gradients) of 32 neighbouring columns share a
// 128-byte line, which is then fetched into one L2 instead of eight.
__device__ __forceinline__ void first_item(const DsgmArgs& A, int item, int* b, bool* pair, int* group) {
    const int nh = (A.H + 3) / 4, nv = (A.W + 3) / 4;
    const int hp = A.B * nh;
    *pair = item < hp;
    const int local = *pair ? item : dev::xcd_item(item - hp, A.B * nv);
    const int per = *pair ? nh : nv;
    *b = local / per;
    *group = local - *b * per;
}

template <int D>
__global__ __launch_bounds__(64) void dsgm_forward_first_kernel(DsgmArgs A) {
    int b, group;
    bool pair;
    first_item(A, blockIdx.x, &b, &pair, &group);
    if (pair) {
        fwd_lines<D, kFwdStoreS>(A, b, 1, group, threadIdx.x);
        fwd_lines<D, kFwdAddS>(A, b, 3, group, threadIdx.x);
    } else {
        fwd_lines<D, kFwdStoreY>(A, b, 0, group, threadIdx.x);
    }
}

template <int D>
__global__ __launch_bounds__(64) void dsgm_forward_last_kernel(DsgmArgs A) {
    const int nv = (A.W + 3) / 4;
    const int item = dev::xcd_item(blockIdx.x, gridDim.x);
    const int b = item / nv, group = item - b * nv;
    fwd_lines<D, kFwdLast>(A, b, 2, group, threadIdx.x);
}

// ------------------------------------------------------------ backward ---
// gS(p, d) = gdisp(p) * (-1/T) * prob(p, d) * (d - disp(p)) of one pixel,
// from S and the forward's minimum mn and normaliser z of that pixel.
template <int DPD>
__device__ __forceinline__ void grad_s_from(const float (&s)[DPD], float mn, float z, float gd,
                                            float dq, float invT, int l, float (&g)[DPD]) {
    const float coef = gd * (-invT) / z;
    const float c = soft_scale(invT);
#pragma unroll
    for (int k = 0; k < DPD; ++k)
        g[k] = coef * soft_weight(s[k], mn, c) * ((float)(l * DPD + k) - dq);
}
template <int DPD>
__device__ __forceinline__ void grad_s_at(const DsgmArgs& A, const float* Sb, const float* stats,
                                          const float* disp, const float* gdisp, int64_t q, int l,
                                          float (&g)[DPD]) {
    float s[DPD];
    loadf<DPD>(Sb + q * A.D, s);
    grad_s_from<DPD>(s, loads(stats + 2 * q), loads(stats + 2 * q + 1), loads(gdisp + q),
                     loads(disp + q), A.inv_temperature, l, g);
}

// What a backward pass does with each pixel's gradients of its direction.
enum BwdMode {
    kBwdStore = 0,    // dC = G, dP1 = d1, dP2 = d2            (left->right)
    kBwdAdd = 1,      // dC += G, dP1 += d1, dP2 += d2         (right->left)
    kBwdStoreY = 2,   // workspace: gY = G, gY1 = d1, gY2 = d2 (top->bottom)
    kBwdLast = 3,     // dC = (dC + gY) + G, likewise dP1, dP2 (bottom->top)
};

// Loads ahead in the backward: what transition s needs (its winner codes
// and k*, S and the soft-argmin statistics of the pixel before it, and, for
// the passes that accumulate, the gradients the pixel already holds) is
// issued kAheadV transitions earlier, into register slots indexed by the
// position within an unrolled block.  Loads for transitions below 1 read
// transition 1 again and are not used.
constexpr int kAheadV = 4;

// One direction's reverse recurrence over a scanline group.
template <int D, int MODE>
__device__ __forceinline__ void bwd_lines(const DsgmArgs& A, int b, int path, int group, int lane) {
    constexpr int DPD = D / 16;
    constexpr int V = kAheadV;
    constexpr bool kPriorC = MODE == kBwdAdd || MODE == kBwdLast;
    constexpr bool kPriorY = MODE == kBwdLast;
    const int sub = lane >> 4, l = lane & 15;
    const LineGeo g = line_geo(A, path, group, sub);
    const int64_t img = (int64_t)A.H * A.W;
    const int64_t npix = (int64_t)A.B * img;
    const float* Sb = A.S + (int64_t)b * img * D + l * DPD;
    const float* stats = A.stats + (int64_t)b * img * 2;
    const float* dsp = A.disp + (int64_t)b * img;
    const float* gdisp = A.grad_disp + (int64_t)b * img;
    const WinWord<D>* win = reinterpret_cast<const WinWord<D>*>(A.wins) +
                            ((int64_t)path * npix + (int64_t)b * img) * 16 + l;
    const uint8_t* kin = A.kstar + (int64_t)path * npix + (int64_t)b * img;
    float* out_c = A.grad_cost + (int64_t)b * img * D + l * DPD;
    float* out_p1 = A.grad_p1 + (int64_t)b * img;
    float* out_p2 = A.grad_p2 + (int64_t)b * img;
    float* gy = A.vol + (int64_t)b * img * D + l * DPD;
    float* gy1 = A.vol_p1 + (int64_t)b * img;
    float* gy2 = A.vol_p2 + (int64_t)b * img;
    const int n = g.n;
    const float invT = A.inv_temperature;
    auto pix_at = [&](int s) __attribute__((always_inline)) { return g.pix0 + (int64_t)s * g.pstep; };

    // inputs of the transitions in flight
    uint32_t vw[V], vk[V];
    float vs[V][DPD], vmn[V], vz[V], vgd[V], vdq[V];
    float vo[V][DPD], vo1[V], vo2[V];
    float vy[V][DPD], vy1[V], vy2[V];
    auto issue = [&](int slot, int s) __attribute__((always_inline)) {
        if (n < 2) return;   // no transitions
        const int64_t q = pix_at(max(s, 1)), qp = q - g.pstep;
        vw[slot] = win[q * 16];
        vk[slot] = kin[q];
        loadf<DPD>(Sb + qp * D, vs[slot]);
        vmn[slot] = loads(stats + 2 * qp);
        vz[slot] = loads(stats + 2 * qp + 1);
        vgd[slot] = loads(gdisp + qp);
        vdq[slot] = loads(dsp + qp);
        if constexpr (kPriorC) {
            loadf<DPD>(out_c + q * D, vo[slot]);
            vo1[slot] = loads(out_p1 + q);
            vo2[slot] = loads(out_p2 + q);
        }
        if constexpr (kPriorY) {
            loadf<DPD>(gy + q * D, vy[slot]);
            vy1[slot] = loads(gy1 + q);
            vy2[slot] = loads(gy2 + q);
        }
    };
    // writes this direction's gradients of one pixel; o, o1, o2 (and y, y1,
    // y2) are what the pixel held before, for the passes that accumulate
    auto emit = [&](int64_t q, const float (&G)[DPD], float d1, float d2, const float (&o)[DPD],
                    float o1, float o2, const float (&y)[DPD], float y1, float y2)
        __attribute__((always_inline)) {
        if (!g.active) return;
        if constexpr (MODE == kBwdStore || MODE == kBwdStoreY) {
            float* oc = MODE == kBwdStore ? out_c : gy;
            storef<DPD>(oc + q * D, G);
            if (l == 0) {
                (MODE == kBwdStore ? out_p1 : gy1)[q] = d1;
                (MODE == kBwdStore ? out_p2 : gy2)[q] = d2;
            }
        } else {
            float v[DPD];
#pragma unroll
            for (int k = 0; k < DPD; ++k) v[k] = (MODE == kBwdLast) ? (o[k] + y[k]) + G[k] : o[k] + G[k];
            storef<DPD>(out_c + q * D, v);
            if (l == 0) {
                out_p1[q] = (MODE == kBwdLast) ? (o1 + y1) + d1 : o1 + d1;
                out_p2[q] = (MODE == kBwdLast) ? (o2 + y2) + d2 : o2 + d2;
            }
        }
    };

    float G[DPD];
    grad_s_at<DPD>(A, Sb, stats, dsp, gdisp, pix_at(n - 1), l, G);
#pragma unroll
    for (int u = 0; u < V; ++u) issue(u, n - 1 - u);
    for (int s0 = n - 1; s0 >= 1; s0 -= V) {
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int s = s0 - u;    // the transition s-1 -> s, in slot u
            if (s >= 1) {
                const int64_t pix = pix_at(s);
                const uint32_t w = vw[u];
                float all = 0.f, s3 = 0.f, s12 = 0.f;
                float A1[DPD], A2[DPD];
#pragma unroll
                for (int k = 0; k < DPD; ++k) {
                    const uint32_t wk = (w >> (2 * k)) & 3u;
                    all += G[k];
                    s3 += (wk == 3u) ? G[k] : 0.f;
                    s12 += (wk == 1u || wk == 2u) ? G[k] : 0.f;
                    A1[k] = (wk == 1u) ? G[k] : 0.f;   // flows to d-1
                    A2[k] = (wk == 2u) ? G[k] : 0.f;   // flows to d+1
                }
                all = row_sumf(all);
                s3 = row_sumf(s3);
                s12 = row_sumf(s12);
                emit(pix, G, s12, s3, vo[u], vo1[u], vo2[u], vy[u], vy1[u], vy2[u]);
                const float a1_next = dppf<kShl1>(0.f, A1[0]);        // lane+1, d = lane end + 1
                const float a2_prev = dppf<kShr1>(0.f, A2[DPD - 1]);  // lane-1, d = lane start - 1
                const int ks = (int)vk[u];
                float gprev[DPD];
                grad_s_from<DPD>(vs[u], vmn[u], vz[u], vgd[u], vdq[u], invT, l, gprev);
                const float jump = s3 - all;
#pragma unroll
                for (int k = 0; k < DPD; ++k) {
                    const uint32_t wk = (w >> (2 * k)) & 3u;
                    float x = gprev[k];
                    x += (wk == 0u) ? G[k] : 0.f;
                    x += (k == DPD - 1) ? a1_next : A1[k + 1];
                    x += (k == 0) ? a2_prev : A2[k - 1];
                    x += (l * DPD + k == ks) ? jump : 0.f;
                    gprev[k] = x;
                }
#pragma unroll
                for (int k = 0; k < DPD; ++k) G[k] = gprev[k];
            }
            issue(u, s - V);
        }
    }
    // the first pixel (no penalty terms)
    float o0[DPD], y0[DPD], o01 = 0.f, o02 = 0.f, y01 = 0.f, y02 = 0.f;
    const int64_t q0 = g.pix0;
    if constexpr (kPriorC) {
        loadf<DPD>(out_c + q0 * D, o0);
        o01 = loads(out_p1 + q0);
        o02 = loads(out_p2 + q0);
    }
    if constexpr (kPriorY) {
        loadf<DPD>(gy + q0 * D, y0);
        y01 = loads(gy1 + q0);
        y02 = loads(gy2 + q0);
    }
    emit(q0, G, 0.f, 0.f, o0, o01, o02, y0, y01, y02);
}

template <int D>
__global__ __launch_bounds__(64) void dsgm_backward_first_kernel(DsgmArgs A) {
    int b, group;
    bool pair;
    first_item(A, blockIdx.x, &b, &pair, &group);
    if (pair) {
        bwd_lines<D, kBwdStore>(A, b, 1, group, threadIdx.x);
        bwd_lines<D, kBwdAdd>(A, b, 3, group, threadIdx.x);
    } else {
        bwd_lines<D, kBwdStoreY>(A, b, 0, group, threadIdx.x);
    }
}

template <int D>
__global__ __launch_bounds__(64) void dsgm_backward_last_kernel(DsgmArgs A) {
    const int nv = (A.W + 3) / 4;
    const int item = dev::xcd_item(blockIdx.x, gridDim.x);
    const int b = item / nv, group = item - b * nv;
    bwd_lines<D, kBwdLast>(A, b, 2, group, threadIdx.x);
}

}  // namespace

#define FSGM_DSGM_DISPATCH(KER, grid, block, A, s)                                       \
    do {                                                                                \
        switch ((A).D) {                                                                \
            case 64: hipLaunchKernelGGL(KER<64>, grid, block, 0, s, A); break;          \
            case 128: hipLaunchKernelGGL(KER<128>, grid, block, 0, s, A); break;        \
            case 256: hipLaunchKernelGGL(KER<256>, grid, block, 0, s, A); break;        \
            default: return hipErrorInvalidValue;                                       \
        }                                                                               \
    } while (0)

// first launch: the horizontal pairs and the top->bottom groups of every
// image; second launch: the bottom->top groups
static unsigned dsgm_items_first(const DsgmArgs& A) {
    return (unsigned)(A.B * ((A.H + 3) / 4 + (A.W + 3) / 4));
}
static unsigned dsgm_items_last(const DsgmArgs& A) { return (unsigned)(A.B * ((A.W + 3) / 4)); }

hipError_t dsgm_launch_forward(const DsgmArgs& A, hipStream_t s) {
    const dim3 b1(64);
    FSGM_DSGM_DISPATCH(dsgm_forward_first_kernel, dim3(dsgm_items_first(A)), b1, A, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FSGM_DSGM_DISPATCH(dsgm_forward_last_kernel, dim3(dsgm_items_last(A)), b1, A, s);
    return hipGetLastError();
}

hipError_t dsgm_launch_backward(const DsgmArgs& A, hipStream_t s) {
    const dim3 b1(64);
    FSGM_DSGM_DISPATCH(dsgm_backward_first_kernel, dim3(dsgm_items_first(A)), b1, A, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FSGM_DSGM_DISPATCH(dsgm_backward_last_kernel, dim3(dsgm_items_last(A)), b1, A, s);
    return hipGetLastError();
}

}  // namespace fsgm
