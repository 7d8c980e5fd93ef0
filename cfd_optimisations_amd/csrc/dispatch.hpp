// dispatch.hpp -- runtime selection of compile-time kernel variants.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

namespace hcfd {

template <int N>
using IC = std::integral_constant<int, N>;

template <typename F>
hipError_t with_width(int w, F&& f) {
    switch (w) {
        case 1: return f(IC<1>{});
        case 8: return f(IC<8>{});
        case 64: return f(IC<64>{});
        default: return hipErrorInvalidValue;
    }
}

template <typename F>
hipError_t with_degree(int d, F&& f) {
    switch (d) {
        case 3: return f(IC<3>{});
        case 4: return f(IC<4>{});
        case 8: return f(IC<8>{});
        default: return hipErrorInvalidValue;
    }
}

template <typename F>
hipError_t with_lanes(int lpc, F&& f) {
    switch (lpc) {
        case 1: return f(IC<1>{});
        case 4: return f(IC<4>{});
        default: return hipErrorInvalidValue;
    }
}

template <typename F>
hipError_t with_bool(bool b, F&& f) {
    if (b) return f(std::true_type{});
    return f(std::false_type{});
}

}  // namespace hcfd
