// Stage-kernel instantiations of the sbp63 family: FP64 single cube, and the dispatch by precision.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o63(const StageLaunch &r) {
    return r.prec == 64 ? launch_stage_f64<63>(r) : launch_stage_o63_f32(r);
}
}  // namespace swm
