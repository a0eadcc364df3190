// Ensemble stage-kernel instantiations of the sbp63 family (FP64 and FP32).
#include "swm_launch.h"
#include "swm_stage.hip"

namespace swm {
hipError_t launch_ens_o63(const EnsLaunch &r) { return r.prec == 64 ? launch_ens_f64<63>(r) : launch_ens_f32<63>(r); }
}  // namespace swm
