// Stage-kernel instantiations of the sbp63 family: FP32 single cube.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o63_f32(const StageLaunch &r) { return launch_stage_f32<63>(r); }
}  // namespace swm
