// MI355X::Flux -- the shape of includes/Base/Flux.hpp:5-19 (MUSCL + Rusanov).
#pragma once
#include "Global.hpp"
#include "MI355X/Domain.hpp"

namespace MI355X {

class Flux {
   public:
    const int nvars;
    const double gma;
    Flux(int nvars, double gma, const Domain& domain);

    // R accumulates into the caller's array, then R *= 1/V (src/Base/Flux.cpp:9-50).
    void ComputeFlux(const double* const U[NVARS], const double* const G[NVARS], const double* const phi[NVARS],
                     double* R[NVARS]) const;

   private:
    const Domain& domain;
};

}  // namespace MI355X
