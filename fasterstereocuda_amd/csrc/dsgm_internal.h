// Internal types of the differentiable SGM kernels (dsgm_kernels.hip).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace fsgm {

struct DsgmArgs {
    const float* cost;         // [B][H][W][D] (the forward's input)
    const float* p1;           // [B][H][W]
    const float* p2;           // [B][H][W]
    float* S;                  // [B][H][W][D]: sum of the four directions (built in place)
    float* disp;               // [B][H][W]
    float* stats;              // [B][H][W][2]: minimum of S and softmax normaliser per pixel
    uint8_t* wins;             // [4][B][H][W][16 lanes] winner codes, D / 64 bytes per lane
    uint8_t* kstar;            // [4][B][H][W]: smallest argmin of the previous step's L
    float* vol;                // [B][H][W][D] workspace: top->bottom L (forward), its dC (backward)
    float* vol_p1;             // [B][H][W] workspace: top->bottom dP1 (backward)
    float* vol_p2;             // [B][H][W] workspace: top->bottom dP2 (backward)
    const float* grad_disp;    // [B][H][W]
    float* grad_cost;          // [B][H][W][D] (built in place)
    float* grad_p1;            // [B][H][W]
    float* grad_p2;            // [B][H][W]
    int32_t B, H, W, D;
    float inv_temperature;
};

// What the forward saves for the backward, in one buffer: the per-pixel
// statistics (2 floats), then the winner codes of every cell of every
// direction (2 bits each: D / 4 bytes per pixel and direction), then each
// pixel's argmin byte per direction.
__host__ __device__ inline int64_t dsgm_saved_bytes(int B, int H, int W, int D) {
    const int64_t npix = (int64_t)B * H * W;
    return npix * 8 + npix * D + npix * 4;
}
__host__ __device__ inline int64_t dsgm_saved_wins_offset(int B, int H, int W) {
    return (int64_t)B * H * W * 8;
}
__host__ __device__ inline int64_t dsgm_saved_kstar_offset(int B, int H, int W, int D) {
    return (int64_t)B * H * W * (8 + D);
}

hipError_t dsgm_launch_forward(const DsgmArgs& A, hipStream_t s);
hipError_t dsgm_launch_backward(const DsgmArgs& A, hipStream_t s);

}  // namespace fsgm
