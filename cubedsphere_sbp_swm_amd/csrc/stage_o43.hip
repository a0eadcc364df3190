// Stage-kernel instantiations of the sbp43 family.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o43(const StageLaunch &r) { return launch_stage_order<43>(r); }
}  // namespace swm
