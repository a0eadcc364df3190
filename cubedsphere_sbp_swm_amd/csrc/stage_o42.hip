// Stage-kernel instantiations of the sbp42 family: FP64 single cube, and the dispatch by precision.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o42(const StageLaunch &r) {
    return r.prec == 64 ? launch_stage_f64<42>(r) : launch_stage_o42_f32(r);
}
}  // namespace swm
