// MI355X boundary conditions -- the shape of includes/Base/BoundaryConditions.hpp:6-65.
// Each object acts on its own face range [F1, F2) on the GPU.
#pragma once
#include "Global.hpp"
#include "MI355X/Domain.hpp"

namespace MI355X {

enum BCtype { SYMMETRY, WALL, SUPINLET, SUBINLET, SUPOUTLET, SUBOUTLET };

class BoundaryCondition {
   public:
    explicit BoundaryCondition(const Domain& domain);
    virtual ~BoundaryCondition() = default;
    void init(int nvars, int F1, int F2, double gma, const double QBC[NVARS]);
    virtual void update_cells(double* U[NVARS]);
    void update_grads(double* G[NVARS]);
    void update_limiter(double* phi[NVARS]);

    int nvars, Nc;
    int F1, F2;
    double gma;
    double QBC[NVARS];
    double UBC[NVARS];
    const Domain& domain;

   protected:
    virtual BCtype type() const = 0;
    int bc = -1;
};

#define HCFD_BC_CLASS(Name, Type)                         \
    class Name : public BoundaryCondition {               \
       public:                                            \
        using BoundaryCondition::BoundaryCondition;       \
                                                          \
       protected:                                         \
        BCtype type() const override { return Type; }     \
    };

HCFD_BC_CLASS(Symmetry, SYMMETRY)
HCFD_BC_CLASS(Wall, WALL)
HCFD_BC_CLASS(SupInlet, SUPINLET)
HCFD_BC_CLASS(SupOutlet, SUPOUTLET)
HCFD_BC_CLASS(SubInlet, SUBINLET)
HCFD_BC_CLASS(SubOutlet, SUBOUTLET)
#undef HCFD_BC_CLASS

}  // namespace MI355X
