// passes_mixed.hip -- FP32-mixed variants of the fused passes.
#include "dispatch.hpp"
#include "passes_impl.hpp"

namespace hcfd {

hipError_t launch_pass_a_mixed(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st) {
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(v.cfl, [&](auto CF) {
                return with_bool(a.ninst > 1, [&](auto MU) {
                    return with_lanes(v.lpc, [&](auto LP) {
                        hipLaunchKernelGGL((k_pass_a<float, decltype(W)::value, decltype(MD)::value,
                                                     decltype(CF)::value, decltype(MU)::value, decltype(LP)::value>),
                                           dim3(grid), dim3(block), 0, st, a);
                        return hipGetLastError();
                    });
                });
            });
        });
    });
}

hipError_t launch_pass_b_mixed(const PassVariant& v, const PassArgs& a, int grid, int block, hipStream_t st) {
    return with_width(v.width, [&](auto W) {
        return with_degree(v.maxdeg, [&](auto MD) {
            return with_bool(a.ninst > 1, [&](auto MU) {
                hipLaunchKernelGGL((k_pass_b<true, decltype(W)::value, decltype(MD)::value, decltype(MU)::value>),
                                   dim3(grid), dim3(block), 0, st, a);
                return hipGetLastError();
            });
        });
    });
}

}  // namespace hcfd
