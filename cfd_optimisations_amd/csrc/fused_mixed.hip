// fused_mixed.hip -- FP32-mixed launchers of the single-kernel step and the CFL
// bound kernel (FP64 states in LDS; FP32 gradient, limiter, face-state
// increments and flux in difference form; compensated residual; FP64 update
// and CFL bound).
#include "dispatch.hpp"
#include "fused_impl.hpp"

namespace hcfd {

hipError_t launch_fused_step_mixed(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st) {
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(v.cfl, [&](auto CF) {
                return with_bool(a.ninst > 1, [&](auto MU) {
                    auto k = k_fused_step<float, decltype(W)::value, decltype(MD)::value, decltype(CF)::value,
                                          decltype(MU)::value>;
                    hipError_t e = hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                       (int)lds);
                    if (e != hipSuccess) return e;
                    hipLaunchKernelGGL(k, dim3(grid), dim3(FUSED_TILE), lds, st, a);
                    return hipGetLastError();
                });
            });
        });
    });
}

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
