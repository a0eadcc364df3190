// Stage-kernel instantiations of the sbp63 family.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o63(const StageLaunch &r) { return launch_stage_order<63>(r); }
hipError_t launch_ens_o63(const EnsLaunch &r) { return launch_ens_order<63>(r); }
}  // namespace swm
