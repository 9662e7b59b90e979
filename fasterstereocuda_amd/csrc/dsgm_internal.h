// Internal types of the differentiable SGM kernels (dsgm_kernels.hip).
#pragma once
#include <stdint.h>
#include <hip/hip_runtime.h>

namespace fsgm {

constexpr int kDsgmSeg = 32;   // steps between checkpoints

struct DsgmArgs {
    const float* cost;         // [B][H][W][D]
    const float* p1;           // [B][H][W]
    const float* p2;           // [B][H][W]
    float* paths;              // [B][4][H][W][D] forward scratch (one volume per direction)
    float* ckpt;               // checkpoints, dsgm_ckpt_floats_per_image() per image
    float* S;                  // [B][H][W][D]
    float* disp;               // [B][H][W]
    const float* grad_disp;    // [B][H][W]
    float* grad_s;             // [B][H][W][D] backward scratch
    float* grad_paths;         // [B][4][H][W][D] backward scratch
    float* grad_p1_paths;      // [B][4][H][W]
    float* grad_p2_paths;      // [B][4][H][W]
    float* grad_cost;          // [B][H][W][D]
    float* grad_p1;            // [B][H][W]
    float* grad_p2;            // [B][H][W]
    int32_t B, H, W, D;
    float inv_temperature;
};

// Checkpoint rows (D floats each) before direction `path` within one image:
// directions 0 and 2 are vertical (W lines of ceil(H/32) checkpoints),
// 1 and 3 horizontal (H lines of ceil(W/32)).
__host__ __device__ inline int64_t dsgm_ckpt_path_offset(int H, int W, int /*D*/, int path) {
    const int64_t v = (int64_t)W * ((H + kDsgmSeg - 1) / kDsgmSeg);
    const int64_t h = (int64_t)H * ((W + kDsgmSeg - 1) / kDsgmSeg);
    switch (path) {
        case 0: return 0;
        case 1: return v;
        case 2: return v + h;
        default: return 2 * v + h;
    }
}
__host__ __device__ inline int64_t dsgm_ckpt_floats_per_image(int H, int W, int D) {
    return dsgm_ckpt_path_offset(H, W, D, 3) * D +
           (int64_t)H * ((W + kDsgmSeg - 1) / kDsgmSeg) * D;
}

hipError_t dsgm_launch_forward(const DsgmArgs& A, hipStream_t s);
hipError_t dsgm_launch_backward(const DsgmArgs& A, hipStream_t s);

}  // namespace fsgm
