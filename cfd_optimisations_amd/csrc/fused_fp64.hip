// fused_fp64.hip -- launchers of the single-kernel step and the CFL bound
// kernel: the dispatch by precision and grid, and the FP64 one-tile-per-block
// variants (persistent-grid variants in fused_fp64_loop.hip, FP32-mixed in
// fused_mixed{,_loop}.hip).
#include "fused_impl.hpp"

namespace hcfd {

extern template hipError_t launch_fused_step_t<double, true>(const PassVariant&, const PassArgs&, int, size_t,
                                                             hipStream_t);
extern template hipError_t launch_fused_step_t<float, false>(const PassVariant&, const PassArgs&, int, size_t,
                                                             hipStream_t);
extern template hipError_t launch_fused_step_t<float, true>(const PassVariant&, const PassArgs&, int, size_t,
                                                            hipStream_t);
template hipError_t launch_fused_step_t<double, false>(const PassVariant&, const PassArgs&, int, size_t, hipStream_t);

hipError_t launch_cfl_bound_mixed(const PassVariant& v, const PassArgs& a, int grid, hipStream_t st);

hipError_t launch_fused_step(const PassVariant& v, const PassArgs& a, int grid, size_t lds, hipStream_t st) {
    if (!pass_variant_supported(v)) return hipErrorInvalidValue;
    const bool loop = fused_grid_loops(a, grid);
    if (v.precision == HCFD_MIXED)
        return loop ? launch_fused_step_t<float, true>(v, a, grid, lds, st)
                    : launch_fused_step_t<float, false>(v, a, grid, lds, st);
    return loop ? launch_fused_step_t<double, true>(v, a, grid, lds, st)
                : launch_fused_step_t<double, false>(v, a, grid, lds, st);
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
