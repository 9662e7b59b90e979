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
//   S = ((L_0 + L_1) + L_2) + L_3
//   disp(p) = sum_d d * softmax_d(-S(p, d) / T)
// Subgradients: the first min candidate in the order [L'(d), L'(d-1)+P1,
// L'(d+1)+P1, m'+P2] wins; m' = L'(k*) with k* the smallest argmin.
//
// Memory: the forward stores the volumes of directions 1, 3 and 0; direction
// 2 runs last and forms S from them and its own L at every pixel.  For the
// backward the forward stores L every 32 steps of every scanline
// (checkpoints).  The backward walks each scanline's 32-step segments from
// the last to the first: it recomputes the segment from its checkpoint,
// keeping the winner codes (2 bits per disparity) and k* of every step in
// registers, then runs the reverse recurrence
//   G(s-1, d) = gS(s-1, d) + G(s, d)[w(s,d)=0] + G(s, d+1)[w(s,d+1)=1]
//             + G(s, d-1)[w(s,d-1)=2] + [d = k*(s-1)] (sum_e G(s,e)[w(s,e)=3] - sum_e G(s,e))
// and emits dC_r(s, d) = G(s, d), dP1_r(s) = sum_d G(s,d)[w in {1,2}],
// dP2_r(s) = sum_d G(s,d)[w = 3].  Directions 1, 3, 0 store theirs;
// direction 2 runs last and adds its own to them at every pixel.
//
// Layout per wave as in the integer aggregation: four 16-lane DPP rows, one
// scanline per row, D/16 disparities per lane.
#include <algorithm>
#include <cmath>

#include "dsgm_internal.h"

#pragma clang fp contract(off)

namespace fsgm {
namespace {

constexpr int kSeg = 32;

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

// Geometry of one scanline of one direction.
struct LineGeo {
    int n;            // steps
    int line;         // row (horizontal) or column (vertical)
    bool active;
    int64_t pix0;     // pixel index of step 0 within the image
    int64_t pstep;    // pixel index increment per step
    int nseg;         // checkpoints on this line
    int64_t ckpt0;    // float offset of this line's first checkpoint (per image)
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
    g.nseg = (g.n + kSeg - 1) / kSeg;
    g.ckpt0 = (dsgm_ckpt_path_offset(H, W, A.D, path) + (int64_t)line * g.nseg) * A.D;
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
// S(p) = ((L_0 + L_1) + L_2) + L_3 and its soft-argmin, for one pixel per
// 16-lane row: s receives S, the row's lane 0 the disparity.
template <int DPD>
__device__ __forceinline__ float soft_argmin(const float (&s)[DPD], int l, float invT) {
    const float mn = row_minf(lane_min(s));
    float z = 0.f, acc = 0.f;
#pragma unroll
    for (int k = 0; k < DPD; ++k) {
        const float e = expf(-(s[k] - mn) * invT);
        z += e;
        acc += (float)(l * DPD + k) * e;
    }
    z = row_sumf(z);
    acc = row_sumf(acc);
    return acc / z;
}

// One direction over a scanline group.  FUSE (direction 2, bottom->top,
// launched after the other three): instead of storing its volume, every
// step forms S from the other three volumes and its own L, in the fixed
// order ((L_0 + L_1) + L_2) + L_3, and writes S and the soft-argmin -- the
// separate sum pass would read all four volumes back.
template <int D, bool FUSE>
__device__ __forceinline__ void fwd_lines(const DsgmArgs& A, int b, int path, int group, int lane) {
    constexpr int DPD = D / 16;
    const int sub = lane >> 4, l = lane & 15;
    const LineGeo g = line_geo(A, path, group, sub);
    const int64_t img = (int64_t)A.H * A.W;
    const float* cost = A.cost + (int64_t)b * img * D + l * DPD;
    const float* p1 = A.p1 + (int64_t)b * img;
    const float* p2 = A.p2 + (int64_t)b * img;
    float* out = A.paths + ((int64_t)b * 4 + path) * img * D + l * DPD;
    const float* vol = A.paths + (int64_t)b * 4 * img * D + l * DPD;   // direction r at r * img * D
    float* S = A.S + (int64_t)b * img * D + l * DPD;
    float* disp = A.disp + (int64_t)b * img;
    float* ck = A.ckpt + (int64_t)b * dsgm_ckpt_floats_per_image(A.H, A.W, D) + g.ckpt0 + l * DPD;
    const float invT = A.inv_temperature;

    // the other three directions at pixel q (FUSE)
    float o0[DPD], o1[DPD], o3[DPD];
    auto load_others = [&](int64_t q) __attribute__((always_inline)) {
        loadf<DPD>(vol + q * D, o0);
        loadf<DPD>(vol + (img + q) * D, o1);
        loadf<DPD>(vol + (3 * img + q) * D, o3);
    };
    auto emit = [&](int64_t q, const float (&L)[DPD]) __attribute__((always_inline)) {
        if constexpr (FUSE) {
            float sv[DPD];
#pragma unroll
            for (int k = 0; k < DPD; ++k) sv[k] = ((o0[k] + o1[k]) + L[k]) + o3[k];
            const float d = soft_argmin<DPD>(sv, l, invT);
            if (g.active) {
                storef<DPD>(S + q * D, sv);
                if (l == 0) disp[q] = d;
            }
        } else {
            if (g.active) storef<DPD>(out + q * D, L);
        }
    };

    float L[DPD];
    loadf<DPD>(cost + g.pix0 * D, L);
    if constexpr (FUSE) load_others(g.pix0);
    emit(g.pix0, L);
    if (g.active) storef<DPD>(ck, L);
    float m = row_minf(lane_min(L));
    int64_t pix = g.pix0;
    float c[DPD];
    if (g.n > 1) loadf<DPD>(cost + (pix + g.pstep) * D, c);
    for (int s = 1; s < g.n; ++s) {
        pix += g.pstep;
        float cur[DPD];
#pragma unroll
        for (int k = 0; k < DPD; ++k) cur[k] = c[k];
        const int64_t nxt = (s + 1 < g.n) ? pix + g.pstep : pix;
        loadf<DPD>(cost + nxt * D, c);
        if constexpr (FUSE) load_others(pix);
        (void)fwd_step<DPD>(L, cur, m, loads(p1 + pix), loads(p2 + pix));
        m = row_minf(lane_min(L));
        emit(pix, L);
        if (g.active && (s & (kSeg - 1)) == 0) storef<DPD>(ck + (int64_t)(s / kSeg) * D, L);
    }
}

// Items of a launch over directions 1, 3, 0 (LAST = false) or over
// direction 2 alone (LAST = true), image by image.
template <bool LAST>
__device__ __forceinline__ void item_of(const DsgmArgs& A, int item, int* b, int* path, int* group) {
    const int nh = (A.H + 3) / 4, nv = (A.W + 3) / 4;
    if constexpr (LAST) {
        *b = item / nv;
        *path = 2;
        *group = item - *b * nv;
        return;
    }
    const int per = 2 * nh + nv;
    *b = item / per;
    const int local = item - *b * per;
    if (local < nh) { *path = 1; *group = local; }
    else if (local < 2 * nh) { *path = 3; *group = local - nh; }
    else { *path = 0; *group = local - 2 * nh; }
}

template <int D, bool FUSE>
__global__ __launch_bounds__(64) void dsgm_forward_kernel(DsgmArgs A) {
    int b, path, group;
    item_of<FUSE>(A, blockIdx.x, &b, &path, &group);
    fwd_lines<D, FUSE>(A, b, path, group, threadIdx.x);
}

// ------------------------------------------------------------ backward ---
// gS(p, d) = gdisp(p) * (-1/T) * prob(p, d) * (d - disp(p))
template <int D>
__global__ __launch_bounds__(256) void dsgm_grad_s_kernel(DsgmArgs A) {
    constexpr int DPD = D / 16;
    const int64_t img = (int64_t)A.H * A.W;
    const int64_t p = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    const bool ok = p < (int64_t)A.B * img;
    const int64_t pc = ok ? p : 0;
    float s[DPD];
    loadf<DPD>(A.S + pc * D + l * DPD, s);
    const float mn = row_minf(lane_min(s));
    const float invT = A.inv_temperature;
    float e[DPD];
    float z = 0.f;
#pragma unroll
    for (int k = 0; k < DPD; ++k) {
        e[k] = expf(-(s[k] - mn) * invT);
        z += e[k];
    }
    z = row_sumf(z);
    const float disp = A.disp[pc];
    const float gd = A.grad_disp[pc];
    float g[DPD];
#pragma unroll
    for (int k = 0; k < DPD; ++k)
        g[k] = gd * (-invT) * (e[k] / z) * ((float)(l * DPD + k) - disp);
    if (ok) storef<DPD>(A.grad_s + pc * D + l * DPD, g);
}

// One direction's reverse recurrence over a scanline group.  FUSE
// (direction 2, launched after the other three): instead of storing its
// gradients, every pixel adds them to the other three directions' in the
// fixed order ((g_0 + g_1) + g_2) + g_3 and writes dC, dP1 and dP2.
template <int D, bool FUSE>
__device__ __forceinline__ void bwd_lines(const DsgmArgs& A, int b, int path, int group, int lane) {
    constexpr int DPD = D / 16;
    const int sub = lane >> 4, l = lane & 15;
    const LineGeo g = line_geo(A, path, group, sub);
    const int64_t img = (int64_t)A.H * A.W;
    const float* cost = A.cost + (int64_t)b * img * D + l * DPD;
    const float* p1 = A.p1 + (int64_t)b * img;
    const float* p2 = A.p2 + (int64_t)b * img;
    const float* gs = A.grad_s + (int64_t)b * img * D + l * DPD;
    const float* ck = A.ckpt + (int64_t)b * dsgm_ckpt_floats_per_image(A.H, A.W, D) + g.ckpt0 + l * DPD;
    float* gc = A.grad_paths + ((int64_t)b * 4 + path) * img * D + l * DPD;
    float* gp1 = A.grad_p1_paths + ((int64_t)b * 4 + path) * img;
    float* gp2 = A.grad_p2_paths + ((int64_t)b * 4 + path) * img;
    const float* gv = A.grad_paths + (int64_t)b * 4 * img * D + l * DPD;   // direction r at r * img * D
    const float* a1v = A.grad_p1_paths + (int64_t)b * 4 * img;
    const float* a2v = A.grad_p2_paths + (int64_t)b * 4 * img;
    float* out_c = A.grad_cost + (int64_t)b * img * D + l * DPD;
    float* out_p1 = A.grad_p1 + (int64_t)b * img;
    float* out_p2 = A.grad_p2 + (int64_t)b * img;
    const int n = g.n;
    // this direction's gradients at pixel q: stored, or summed with the others
    auto emit = [&](int64_t q, const float (&G)[DPD], float d1, float d2) __attribute__((always_inline)) {
        if (!g.active) return;
        if constexpr (FUSE) {
            float o0[DPD], o1[DPD], o3[DPD], sv[DPD];
            loadf<DPD>(gv + q * D, o0);
            loadf<DPD>(gv + (img + q) * D, o1);
            loadf<DPD>(gv + (3 * img + q) * D, o3);
#pragma unroll
            for (int k = 0; k < DPD; ++k) sv[k] = ((o0[k] + o1[k]) + G[k]) + o3[k];
            storef<DPD>(out_c + q * D, sv);
            if (l == 0) {
                out_p1[q] = ((a1v[q] + a1v[img + q]) + d1) + a1v[3 * img + q];
                out_p2[q] = ((a2v[q] + a2v[img + q]) + d2) + a2v[3 * img + q];
            }
        } else {
            storef<DPD>(gc + q * D, G);
            if (l == 0) { gp1[q] = d1; gp2[q] = d2; }
        }
    };

    float G[DPD];
    loadf<DPD>(gs + (g.pix0 + (int64_t)(n - 1) * g.pstep) * D, G);
    for (int j = g.nseg - 1; j >= 0; --j) {
        const int s0 = j * kSeg;
        const int s1 = min(s0 + kSeg, n - 1);     // last transition of this segment
        // recompute L(s0 .. s1-1), recording the transitions s0+1 .. s1
        uint32_t wins[kSeg];
        uint32_t kst[kSeg / 4];
#pragma unroll
        for (int t = 0; t < kSeg / 4; ++t) kst[t] = 0u;
        float L[DPD];
        loadf<DPD>(ck + (int64_t)j * D, L);
        float m = row_minf(lane_min(L));
#pragma unroll
        for (int t = 0; t < kSeg; ++t) {
            const int s = s0 + 1 + t;
            if (s <= s1) {
                const int ks = row_argmin<DPD>(L, m, l);
                kst[t >> 2] |= (uint32_t)ks << (8 * (t & 3));
                const int64_t pix = g.pix0 + (int64_t)s * g.pstep;
                float c[DPD];
                loadf<DPD>(cost + pix * D, c);
                wins[t] = fwd_step<DPD>(L, c, m, loads(p1 + pix), loads(p2 + pix));
                m = row_minf(lane_min(L));
            } else {
                wins[t] = 0u;
            }
        }
        // reverse sweep over the transitions s1 .. s0+1
#pragma unroll
        for (int t = kSeg - 1; t >= 0; --t) {
            const int s = s0 + 1 + t;
            if (s <= s1) {
                const int64_t pix = g.pix0 + (int64_t)s * g.pstep;
                const uint32_t w = wins[t];
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
                emit(pix, G, s12, s3);
                const float a1_next = dppf<kShl1>(0.f, A1[0]);        // lane+1, d = lane end + 1
                const float a2_prev = dppf<kShr1>(0.f, A2[DPD - 1]);  // lane-1, d = lane start - 1
                const int ks = (int)((kst[t >> 2] >> (8 * (t & 3))) & 0xFFu);
                float gprev[DPD];
                loadf<DPD>(gs + (pix - g.pstep) * D, gprev);
                const float jump = s3 - all;
#pragma unroll
                for (int k = 0; k < DPD; ++k) {
                    const uint32_t wk = (w >> (2 * k)) & 3u;
                    float v = gprev[k];
                    v += (wk == 0u) ? G[k] : 0.f;
                    v += (k == DPD - 1) ? a1_next : A1[k + 1];
                    v += (k == 0) ? a2_prev : A2[k - 1];
                    v += (l * DPD + k == ks) ? jump : 0.f;
                    gprev[k] = v;
                }
#pragma unroll
                for (int k = 0; k < DPD; ++k) G[k] = gprev[k];
            }
        }
    }
    emit(g.pix0, G, 0.f, 0.f);   // no penalty at the first pixel
}

template <int D, bool FUSE>
__global__ __launch_bounds__(64) void dsgm_backward_kernel(DsgmArgs A) {
    int b, path, group;
    item_of<FUSE>(A, blockIdx.x, &b, &path, &group);
    bwd_lines<D, FUSE>(A, b, path, group, threadIdx.x);
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
#define FSGM_DSGM_DISPATCH2(KER, FUSE, grid, block, A, s)                                \
    do {                                                                                \
        switch ((A).D) {                                                                \
            case 64: hipLaunchKernelGGL((KER<64, FUSE>), grid, block, 0, s, A); break;  \
            case 128: hipLaunchKernelGGL((KER<128, FUSE>), grid, block, 0, s, A); break; \
            case 256: hipLaunchKernelGGL((KER<256, FUSE>), grid, block, 0, s, A); break; \
            default: return hipErrorInvalidValue;                                       \
        }                                                                               \
    } while (0)

// scanline groups of directions 1, 3 and 0, and of direction 2
static int dsgm_items3(const DsgmArgs& A) { return A.B * (2 * ((A.H + 3) / 4) + (A.W + 3) / 4); }
static int dsgm_items1(const DsgmArgs& A) { return A.B * ((A.W + 3) / 4); }

// directions 1, 3, 0 store their volumes; direction 2 then sums the four and
// takes the soft-argmin
hipError_t dsgm_launch_forward(const DsgmArgs& A, hipStream_t s) {
    const dim3 b1(64);
    FSGM_DSGM_DISPATCH2(dsgm_forward_kernel, false, dim3(dsgm_items3(A)), b1, A, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    FSGM_DSGM_DISPATCH2(dsgm_forward_kernel, true, dim3(dsgm_items1(A)), b1, A, s);
    return hipGetLastError();
}

hipError_t dsgm_launch_backward(const DsgmArgs& A, hipStream_t s) {
    const int64_t npix = (int64_t)A.B * A.H * A.W;
    const dim3 g0((unsigned)((npix + 15) / 16)), b0(256);
    FSGM_DSGM_DISPATCH(dsgm_grad_s_kernel, g0, b0, A, s);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 b1(64);
    FSGM_DSGM_DISPATCH2(dsgm_backward_kernel, false, dim3(dsgm_items3(A)), b1, A, s);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    FSGM_DSGM_DISPATCH2(dsgm_backward_kernel, true, dim3(dsgm_items1(A)), b1, A, s);
    return hipGetLastError();
}

}  // namespace fsgm
