// Stage-kernel instantiations of the sbp43 family.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o43(const StageLaunch &r) { return launch_stage_order<43>(r); }
hipError_t launch_ens_o43(const EnsLaunch &r) { return launch_ens_order<43>(r); }
}  // namespace swm
