// fused_fp64.hip -- launchers of the single-kernel step and the CFL bound
// kernel (FP64 variants here; FP32-mixed in fused_mixed.hip).
#include "dispatch.hpp"
#include "fused_impl.hpp"

namespace hcfd {

hipError_t launch_fused_step_mixed(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st);
hipError_t launch_cfl_bound_mixed(const PassVariant& v, const PassArgs& a, int grid, hipStream_t st);

hipError_t launch_fused_step(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st) {
    if (!pass_variant_supported(v)) return hipErrorInvalidValue;
    if (v.precision == HCFD_MIXED) return launch_fused_step_mixed(v, a, grid, lds, st);
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(v.cfl, [&](auto CF) {
                return with_bool(a.ninst > 1, [&](auto MU) {
                    auto k = k_fused_step<double, decltype(W)::value, decltype(MD)::value, decltype(CF)::value,
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

hipError_t launch_cfl_bound(const PassVariant& v, const PassArgs& a, int grid, hipStream_t st) {
    if (!pass_variant_supported(v)) return hipErrorInvalidValue;
    if (v.precision == HCFD_MIXED) return launch_cfl_bound_mixed(v, a, grid, st);
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(a.ninst > 1, [&](auto MU) {
                hipLaunchKernelGGL((k_cfl_bound<double, decltype(W)::value, decltype(MD)::value, decltype(MU)::value>),
                                   dim3(grid), dim3(FUSED_TILE), 0, st, a);
                return hipGetLastError();
            });
        });
    });
}

}  // namespace hcfd
