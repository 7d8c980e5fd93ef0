// MI355X::Limiter -- the shape of includes/Base/Limiters.hpp:6-18.
#pragma once
#include "Global.hpp"
#include "MI355X/Domain.hpp"

namespace MI355X {

class Limiter {
   public:
    explicit Limiter(const Domain& domain);
    void limit_gradient(const int nvars, const double* const U[NVARS], const double* const G[NVARS],
                        double* phi[NVARS]);

   private:
    const Domain& domain;
};

}  // namespace MI355X
