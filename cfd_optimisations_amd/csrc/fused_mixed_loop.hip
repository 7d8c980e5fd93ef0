// fused_mixed_loop.hip -- the FP32-mixed single-kernel step on a persistent
// grid (blocks walk several tiles); dispatched from fused_fp64.hip.
#include "fused_impl.hpp"

namespace hcfd {

template hipError_t launch_fused_step_t<float, true>(const PassVariant&, const PassArgs&, int, size_t, hipStream_t);

}  // namespace hcfd
