// Stage-kernel instantiations of the sbp43 family: FP64 single cube, and the dispatch by precision.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o43(const StageLaunch &r) {
    return r.prec == 64 ? launch_stage_f64<43>(r) : launch_stage_o43_f32(r);
}
}  // namespace swm
