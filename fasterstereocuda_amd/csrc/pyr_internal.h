// Internal types of the coarse-to-fine pyramid kernels (pyr_kernels.hip).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "sgm_internal.h"

namespace fsgm {

struct PyrZoom {               // both views of one level -> the next coarser one
    const uint8_t* src_l;
    const uint8_t* src_r;
    uint8_t* dst_l;
    uint8_t* dst_r;
    int32_t w, h;              // source size
};

struct PyrInit {               // coarse disparity -> base disparity of the finer level
    const float* coarse;       // (w/2) x (h/2)
    int16_t* next;             // w x h
    int32_t w, h, level_range;
};

struct PyrLevel {              // one finer level (left view only)
    const uint64_t* census_l;
    const uint64_t* census_r;
    const int16_t* init;       // w x h base disparities, -32768 = INVALID
    uint8_t* cost;             // [h][w][LR]
    uint8_t* path[4];          // [h][w][LR]
    float* disp;               // w x h WTA output
    int32_t w, h, level_range; // level_range: 16 or 32
    int32_t cs_type, p1, p2;
    float uniqueness;
};

hipError_t pyr_launch_zoom(const PyrZoom& z, hipStream_t s);
hipError_t pyr_launch_init_next(const PyrInit& a, hipStream_t s);
hipError_t pyr_launch_dilation(const float* in, float* out, int w, int h, int wnd, hipStream_t s);
hipError_t pyr_launch_cost_init(const PyrLevel& a, hipStream_t s);
hipError_t pyr_launch_aggregate_init(const PyrLevel& a, hipStream_t s);
// process-wide SGM_WINDOW_* form of the window aggregation
int window_form();
void set_window_form(int form);
hipError_t pyr_launch_wta_init(const PyrLevel& a, hipStream_t s);

// plan upload shared by the plain, batched and pyramid paths (sgm_capi.hip)
int upload_plan(const std::vector<SgmJob>& jobs, const std::vector<SgmPost>& posts,
                StagePlan* plan, void** block, std::string* err);
int set_global_error(int code, const std::string& msg);

}  // namespace fsgm
