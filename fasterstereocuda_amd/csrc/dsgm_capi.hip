// C ABI of the differentiable SGM (declared in sgm_capi.h).  No device
// memory is allocated here: the caller passes every buffer.
#include <string>

#include "dsgm_internal.h"
#include "sgm_capi.h"

using namespace fsgm;

namespace {

bool dsgm_shape_ok(int32_t B, int32_t H, int32_t W, int32_t D) {
    return B > 0 && H > 0 && W > 0 && (D == 64 || D == 128 || D == 256);
}

int dsgm_fail(const char* msg);

}  // namespace

// error text shared with the rest of the C ABI
extern "C" const char* sgm_last_error(const sgm_handle* h);
namespace fsgm { int set_global_error(int code, const std::string& msg); }

namespace {
int dsgm_fail(const char* msg) { return fsgm::set_global_error(SGM_E_INVALID_ARG, msg); }
}

extern "C" {

// saved by the forward for the backward, in floats: the per-pixel minimum of
// S and softmax normaliser (2 floats per pixel), the winner codes (1 byte
// per cell over the four directions) and the argmin bytes (4 per pixel)
size_t sgm_dsgm_checkpoint_floats(int32_t B, int32_t H, int32_t W, int32_t D) {
    if (!dsgm_shape_ok(B, H, W, D)) return 0;
    return (size_t)((dsgm_saved_bytes(B, H, W, D) + 3) / 4);
}

// transient: one direction volume and two per-pixel planes
size_t sgm_dsgm_workspace_floats(int32_t B, int32_t H, int32_t W, int32_t D) {
    if (!dsgm_shape_ok(B, H, W, D)) return 0;
    const size_t npix = (size_t)B * H * W;
    return npix * D + 2 * npix;
}

int sgm_dsgm_forward(const float* d_cost, const float* d_p1, const float* d_p2, int32_t B,
                     int32_t H, int32_t W, int32_t D, float temperature, float* d_S,
                     float* d_disp, float* d_ckpt, float* d_work, void* stream) {
    if (!dsgm_shape_ok(B, H, W, D)) return dsgm_fail("dsgm: bad shape (D must be 64, 128 or 256)");
    if (!(temperature > 0.f)) return dsgm_fail("dsgm: temperature must be > 0");
    if (!d_cost || !d_p1 || !d_p2 || !d_S || !d_disp || !d_ckpt || !d_work)
        return dsgm_fail("dsgm: NULL pointer");
    DsgmArgs A{};
    A.cost = d_cost; A.p1 = d_p1; A.p2 = d_p2;
    A.vol = d_work;
    A.S = d_S; A.disp = d_disp;
    A.stats = d_ckpt;
    A.wins = reinterpret_cast<uint8_t*>(d_ckpt) + dsgm_saved_wins_offset(B, H, W);
    A.kstar = reinterpret_cast<uint8_t*>(d_ckpt) + dsgm_saved_kstar_offset(B, H, W, D);
    A.B = B; A.H = H; A.W = W; A.D = D;
    A.inv_temperature = 1.0f / temperature;
    const hipError_t e = dsgm_launch_forward(A, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fsgm::set_global_error(SGM_E_HIP, hipGetErrorString(e));
    return SGM_OK;
}

int sgm_dsgm_backward(const float* d_cost, const float* d_p1, const float* d_p2,
                      const float* d_S, const float* d_disp, const float* d_ckpt,
                      const float* d_grad_disp, int32_t B, int32_t H, int32_t W, int32_t D,
                      float temperature, float* d_grad_cost, float* d_grad_p1,
                      float* d_grad_p2, float* d_work, void* stream) {
    if (!dsgm_shape_ok(B, H, W, D)) return dsgm_fail("dsgm: bad shape (D must be 64, 128 or 256)");
    if (!(temperature > 0.f)) return dsgm_fail("dsgm: temperature must be > 0");
    if (!d_cost || !d_p1 || !d_p2 || !d_S || !d_disp || !d_ckpt || !d_grad_disp ||
        !d_grad_cost || !d_grad_p1 || !d_grad_p2 || !d_work)
        return dsgm_fail("dsgm: NULL pointer");
    const size_t npix = (size_t)B * H * W;
    DsgmArgs A{};
    A.cost = d_cost; A.p1 = d_p1; A.p2 = d_p2;
    A.stats = const_cast<float*>(d_ckpt);
    A.wins = reinterpret_cast<uint8_t*>(const_cast<float*>(d_ckpt)) + dsgm_saved_wins_offset(B, H, W);
    A.kstar = reinterpret_cast<uint8_t*>(const_cast<float*>(d_ckpt)) +
              dsgm_saved_kstar_offset(B, H, W, D);
    A.S = const_cast<float*>(d_S);
    A.disp = const_cast<float*>(d_disp);
    A.grad_disp = d_grad_disp;
    A.vol = d_work;
    A.vol_p1 = d_work + npix * D;
    A.vol_p2 = d_work + npix * D + npix;
    A.grad_cost = d_grad_cost; A.grad_p1 = d_grad_p1; A.grad_p2 = d_grad_p2;
    A.B = B; A.H = H; A.W = W; A.D = D;
    A.inv_temperature = 1.0f / temperature;
    const hipError_t e = dsgm_launch_backward(A, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fsgm::set_global_error(SGM_E_HIP, hipGetErrorString(e));
    return SGM_OK;
}

}  // extern "C"
