// Launch interface between the C ABI (swm_capi.hip) and the per-order kernel translation units
// (stage_o42.hip, ...): one entry point per SBP family, precision and tile height chosen at run
// time.
#pragma once

#include <hip/hip_runtime.h>

#include "swm_kernels.h"

namespace swm {

struct StageLaunch {
    int prec;   // 64 or 32
    int mode;   // StageMode
    int ty;     // tile height (8 or 16); tile width is 64
    bool otf;   // FP64: metric terms evaluated on the fly
    hipStream_t stream;
    DevGeo<double> gd;
    DevGeo<float> gf;
    StageArgs<double> ad;
    StageArgs<float> af;
};

// ensemble stage launch: member table (EnsMember<T>[nmembers]) and tile prefix sums on the
// device; the persistent grid is num_cu x resident blocks per CU x blocks_per_slot
struct EnsLaunch {
    int prec, mode, ty;
    bool otf;
    hipStream_t stream;
    const void *members;
    const int *tile_start;
    int nmembers, ntiles, num_cu, blocks_per_slot;
};

hipError_t launch_ens_o42(const EnsLaunch &r);
hipError_t launch_ens_o43(const EnsLaunch &r);
hipError_t launch_ens_o63(const EnsLaunch &r);

hipError_t launch_stage_o21(const StageLaunch &r);
hipError_t launch_stage_o42(const StageLaunch &r);
hipError_t launch_stage_o43(const StageLaunch &r);
hipError_t launch_stage_o63(const StageLaunch &r);

}  // namespace swm
