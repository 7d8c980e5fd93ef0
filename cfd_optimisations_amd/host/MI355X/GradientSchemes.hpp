// MI355X gradient schemes -- the shape of includes/Base/GradientSchemes.hpp:6-29.
#pragma once
#include "Global.hpp"
#include "MI355X/Domain.hpp"

namespace MI355X {

class GradientScheme {
   public:
    explicit GradientScheme(const Domain& domain);
    virtual ~GradientScheme() = default;
    virtual void compute_gradient(int nvars, const double* const u[NVARS], double* grad_u[NVARS]) const;

    const int Nc, Nf;
    const Domain& domain;
};

// Green-Gauss on the GPU (GaussGreen::compute_gradient, src/Base/GaussGreen.cpp:14-40).
class GaussGreen : public GradientScheme {
   public:
    using GradientScheme::GradientScheme;
    void compute_gradient(int nvars, const double* const u[NVARS], double* grad_u[NVARS]) const override;
};

class NoGrads : public GradientScheme {
   public:
    using GradientScheme::GradientScheme;
};

}  // namespace MI355X
