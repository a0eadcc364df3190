// Stage-kernel instantiations of the sbp42 family.
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_stage_o42(const StageLaunch &r) { return launch_stage_order<42>(r); }
hipError_t launch_ens_o42(const EnsLaunch &r) { return launch_ens_order<42>(r); }
}  // namespace swm
