// Peak removal on the GPU: small connected segments of a disparity map are
// set invalid.
//
// The reference (SgmStereoCuda/StereoCudaImpl.cpp:32-97, 463-498) runs, on
// the host after the map has been copied back, a breadth-first search per
// horizontal strip: strip i of `strips` covers rows [i r, (i + 1) r) with
// r = H / strips (rows from strips * r on are left as they are); within a
// strip, valid 4-neighbours join one segment when |d - d'| <= 1 (float
// arithmetic), and a segment of fewer than (int)((float)(W r) * ratio)
// pixels becomes invalid.  The reference takes `strips` from the host's core
// count (hardware_concurrency() * 3 / 4), which makes its result depend on
// the machine; here it is a parameter (default 1, the whole image).
//
// The segments are the connected components of the symmetric relation
// above, so any method that finds components gives the same answer.  This
// one is a lock-free union-find over the right and down edges: parents only
// ever decrease (parent[x] <= x), every union retry strictly lowers one of
// its two indices and every find strictly lowers its index, so all loops end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <mutex>

#include "sgm_internal.h"

namespace fsgm {
namespace {

// relaxed read of a parent link another lane may be lowering
__device__ __forceinline__ int parent_of(const int* p, int x) {
    return __hip_atomic_load(p + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int find_root(const int* p, int x) {
    int q = parent_of(p, x);
    while (q != x) {
        x = q;
        q = parent_of(p, x);
    }
    return x;
}

__device__ __forceinline__ void unite(int* p, int a, int b) {
    while (true) {
        a = find_root(p, a);
        b = find_root(p, b);
        if (a == b) return;
        if (a > b) { const int t = a; a = b; b = t; }
        const int old = atomicMin(p + b, a);   // link root b under a
        if (old == b) return;                  // b was still a root
        b = old;                               // b moved meanwhile: retry from its new parent
    }
}

__device__ __forceinline__ bool joins(float a, float b) {
    return a != kInvalidDisp && b != kInvalidDisp && fabsf(a - b) <= 1.0f;
}

// parent[i] = i, size[i] = 0 over the processed rows
__global__ __launch_bounds__(256) void peaks_init_kernel(int* __restrict__ parent,
                                                         int* __restrict__ size, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        parent[i] = i;
        size[i] = 0;
    }
}

// union over the right and down edges inside one strip
__global__ __launch_bounds__(256) void peaks_union_kernel(const float* __restrict__ disp,
                                                          int* __restrict__ parent, int W,
                                                          int rows, int strip_rows) {
    const int n = W * rows;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float d = disp[i];
        if (d == kInvalidDisp) continue;
        const int x = i % W, y = i / W;
        if (x + 1 < W && joins(d, disp[i + 1])) unite(parent, i, i + 1);
        if (y + 1 < rows && (y + 1) % strip_rows != 0 && joins(d, disp[i + W]))
            unite(parent, i, i + W);
    }
}

// segment sizes, counted at the roots
__global__ __launch_bounds__(256) void peaks_count_kernel(const float* __restrict__ disp,
                                                          int* __restrict__ parent,
                                                          int* __restrict__ size, int n) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (disp[i] == kInvalidDisp) continue;
        const int r = find_root(parent, i);
        atomicAdd(size + r, 1);
    }
}

__global__ __launch_bounds__(256) void peaks_remove_kernel(float* __restrict__ disp,
                                                           const int* __restrict__ parent,
                                                           const int* __restrict__ size, int n,
                                                           int threshold) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (disp[i] == kInvalidDisp) continue;
        if (size[find_root(parent, i)] < threshold) disp[i] = kInvalidDisp;
    }
}

}  // namespace

int peaks_threshold(int W, int strip_rows, float ratio) {
    // StereoCudaImpl.cpp:476: int * int, times a float, truncated to int
    return (int)((float)(W * strip_rows) * ratio);
}

// Scratch of the allocation-free peak removal (sgm_stage_remove_peaks_noalloc):
// parent links and segment sizes of up to kPeaksNoAllocMaxPixels pixels,
// 2 * 4 B per pixel (16 MiB).  Each device holds its own copy for the life of
// the process.  Host code never uses this name as an address: it passes the
// per-device pointer from hipGetSymbolAddress.
static __device__ __align__(256) int scratch[2 * (size_t)kPeaksNoAllocMaxPixels];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t peaks_noalloc_scratch(int** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        // once per device, with that device current: its copy's address
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<int*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

hipError_t launch_remove_peaks(float* disp, int* work, int W, int H, int strips, float ratio,
                               hipStream_t s) {
    if (W <= 0 || H <= 0 || strips <= 0) return hipErrorInvalidValue;
    const int strip_rows = H / strips;
    const int rows = strip_rows * strips;   // rows past the last strip are left alone
    if (strip_rows == 0) return hipSuccess;
    const int threshold = peaks_threshold(W, strip_rows, ratio);
    const int n = W * rows;
    int* parent = work;
    int* size = work + (size_t)W * H;
    const unsigned blocks = (unsigned)std::min((n + 255) / 256, 65535);
    hipLaunchKernelGGL(peaks_init_kernel, dim3(blocks), dim3(256), 0, s, parent, size, n);
    hipLaunchKernelGGL(peaks_union_kernel, dim3(blocks), dim3(256), 0, s, disp, parent, W, rows,
                       strip_rows);
    hipLaunchKernelGGL(peaks_count_kernel, dim3(blocks), dim3(256), 0, s, disp, parent, size, n);
    hipLaunchKernelGGL(peaks_remove_kernel, dim3(blocks), dim3(256), 0, s, disp, parent, size, n,
                       threshold);
    return hipGetLastError();
}

}  // namespace fsgm
