// Stage-kernel instantiations of the sbp43 family: FP32 single cube.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o43_f32(const StageLaunch &r) { return launch_stage_f32<43>(r); }
}  // namespace swm
