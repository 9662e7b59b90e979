// Device helpers shared by the SGM kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fsgm {
namespace dev {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_v(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) { return as_u(as_v(a) + as_v(b)); }
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) { return as_u(as_v(a) - as_v(b)); }
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return as_u(__builtin_elementwise_min(as_v(a), as_v(b)));
}

// byte <-> packed-u16 conversions (v_perm_b32)
__device__ __forceinline__ uint32_t bytes01(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0C010C00u); }
__device__ __forceinline__ uint32_t bytes23(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0C030C02u); }
// low byte of each u16 of lo and hi -> [lo.0, lo.1, hi.0, hi.1] (uint8 truncation)
__device__ __forceinline__ uint32_t pack4(uint32_t lo, uint32_t hi) {
    return __builtin_amdgcn_perm(hi, lo, 0x06040200u);
}

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)v, CTRL, 0xF, 0xF, false);
}
// The same move where no lane keeps an old value: a lane whose source lies
// outside its DPP row reads 0 (bound_ctrl).  For patterns whose every source
// is inside the row (quad_perm, mirrors), or whose out-of-row lanes are
// overwritten afterwards, it saves the copy of `old` into the destination.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp0(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
constexpr int kRowShl1 = 0x101;
constexpr int kRowShr1 = 0x111;

// minimum over the 16 lanes of a DPP row; every lane receives it
__device__ __forceinline__ uint32_t row_min16(uint32_t m) {
    m = min(m, dpp<0x128>(0u, m));
    m = min(m, dpp<0x124>(0u, m));
    m = min(m, dpp<0x122>(0u, m));
    m = min(m, dpp<0x121>(0u, m));
    return m;
}

__device__ __forceinline__ int find_job(const int* __restrict__ offs, int n, int item) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= item) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Workgroups of one launch go round-robin to the 8 XCDs (block b to XCD
// b mod 8), and each XCD has its own L2.  This bijection on [0, n) gives each
// XCD one contiguous run of items, so neighbouring items (adjacent column
// groups that read overlapping census rows) run at the same time under one
// L2.  On a part with another XCD count it is still a bijection.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item(int b, int n) {
    const int q = n / kXcds, r = n % kXcds;
    const int x = b % kXcds, i = b / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + i;
}

// Global-address-space views so loads/stores compile to global_* (not flat_*).
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) u32x2 g_u2;
typedef __attribute__((address_space(1))) u32x4 g_u4;

// Scalar element access through the global address space.  Generic (flat)
// accesses count on both the vector-memory and the LDS counters, which makes
// the compiler drain every outstanding load before the next use of any.
template <class T>
__device__ __forceinline__ T gld(const T* p) {
    return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
    *(__attribute__((address_space(1))) T*)p = v;
}

template <int NW>
__device__ __forceinline__ void load_words(const uint8_t* p, uint32_t (&w)[NW]) {
    if constexpr (NW == 1) {
        w[0] = *(const g_u32*)p;
    } else if constexpr (NW == 2) {
        const u32x2 v = *(const g_u2*)p;
        w[0] = v.x; w[1] = v.y;
    } else {
#pragma unroll
        for (int k = 0; k < NW; k += 4) {
            const u32x4 v = *(const g_u4*)(p + 4 * k);
            w[k] = v.x; w[k + 1] = v.y; w[k + 2] = v.z; w[k + 3] = v.w;
        }
    }
}

template <int NW>
__device__ __forceinline__ void store_words(uint8_t* p, const uint32_t (&w)[NW]) {
    if constexpr (NW == 1) {
        *(g_u32*)p = w[0];
    } else if constexpr (NW == 2) {
        *(g_u2*)p = u32x2{w[0], w[1]};
    } else {
#pragma unroll
        for (int k = 0; k < NW; k += 4)
            *(g_u4*)(p + 4 * k) = u32x4{w[k], w[k + 1], w[k + 2], w[k + 3]};
    }
}

}  // namespace dev
}  // namespace fsgm
