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
    int tx, ty; // tile shape (tile_shape_ok); the sbp21 kernel uses 64 x 16 for ty = 16, else 64 x 8
    bool otf;   // FP64: metric terms evaluated on the fly
    hipStream_t stream;
    DevGeo<double> gd;
    DevGeo<float> gf;
    StageArgs<double> ad;
    StageArgs<float> af;
};

// ensemble stage launch: member table (EnsMember<T>[nmembers]) and task table (member, task
// index within the member) on the device; the grid is num_cu x resident blocks per CU x
// blocks_per_slot (a persistent grid), or one block per task for blocks_per_slot = 0
struct EnsLaunch {
    int prec, mode, tx, ty;
    bool otf;
    hipStream_t stream;
    const void *members;
    const int2 *tasks;
    int nmembers, ntiles, num_cu, blocks_per_slot;
};

// one translation unit per (family, precision / ensemble), compiled in parallel:
// stage_oNN.hip (FP64 single cube, and the dispatch), stage_oNN_f32.hip, ens_oNN.hip
hipError_t launch_ens_o42(const EnsLaunch &r);
hipError_t launch_ens_o43(const EnsLaunch &r);
hipError_t launch_ens_o63(const EnsLaunch &r);

hipError_t launch_stage_o21(const StageLaunch &r);
hipError_t launch_stage_o42(const StageLaunch &r);
hipError_t launch_stage_o43(const StageLaunch &r);
hipError_t launch_stage_o63(const StageLaunch &r);
hipError_t launch_stage_o42_f32(const StageLaunch &r);
hipError_t launch_stage_o43_f32(const StageLaunch &r);
hipError_t launch_stage_o63_f32(const StageLaunch &r);

}  // namespace swm
