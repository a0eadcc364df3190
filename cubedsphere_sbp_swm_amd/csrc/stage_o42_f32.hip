// Stage-kernel instantiations of the sbp42 family: FP32 single cube.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o42_f32(const StageLaunch &r) { return launch_stage_f32<42>(r); }
}  // namespace swm
