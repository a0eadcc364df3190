// Ensemble stage-kernel instantiations of the sbp43 family (FP64 and FP32).
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_ens_o43(const EnsLaunch &r) { return r.prec == 64 ? launch_ens_f64<43>(r) : launch_ens_f32<43>(r); }
}  // namespace swm
