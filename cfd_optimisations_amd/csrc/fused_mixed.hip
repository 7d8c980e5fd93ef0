// fused_mixed.hip -- FP32-mixed single-kernel step, one tile per block, and
// the mixed CFL bound kernel (FP64 states in LDS; FP32 gradient, limiter,
// face-state increments and flux in difference form; FP64 residual sum,
// update and CFL bound).  Dispatched from fused_fp64.hip.
#include "fused_impl.hpp"

namespace hcfd {

template hipError_t launch_fused_step_t<float, false>(const PassVariant&, const PassArgs&, int, size_t, hipStream_t);

hipError_t launch_cfl_bound_mixed(const PassVariant& v, const PassArgs& a, int grid, hipStream_t st) {
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(a.ninst > 1, [&](auto MU) {
                hipLaunchKernelGGL((k_cfl_bound<float, decltype(W)::value, decltype(MD)::value, decltype(MU)::value>),
                                   dim3(grid), dim3(FUSED_TILE), 0, st, a);
                return hipGetLastError();
            });
        });
    });
}

}  // namespace hcfd
