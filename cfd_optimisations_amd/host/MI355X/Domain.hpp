// MI355X::Domain -- the shape of Base::Domain (includes/Base/Domain.hpp:6-38).
// Construction builds the device context (CSR, RCM, colours, geometry on the
// GPU); the public geometry arrays are host copies in the mesh's order.
#pragma once
#include <vector>

#include "Global.hpp"
#include "Mesh.hpp"
#include "hcfd.h"

namespace MI355X {

class Domain {
   public:
    const int Nf, Nc;
    const int Ncol, NBC;

    Mesh& mesh;

    int (*C)[2];
    double (*rf)[2][NDIMS];
    double (*nf)[NDIMS];
    double (*xf)[NDIMS];
    double* Af;
    double* If;

    double (*xc)[NDIMS];
    double* iv;

    // precision: HCFD_FP64 or HCFD_MIXED; flags: hcfd_flags.  The FP64 step is
    // bit-identical to Base in every cell order (HCFD_FLAG_EXACT only restates
    // that, and is refused with HCFD_MIXED); the default Hilbert order gives the
    // compact tiles the single-kernel step needs (flags 0 keep reverse
    // Cuthill-McKee).
    explicit Domain(Mesh& mesh, int precision = HCFD_FP64, int flags = HCFD_FLAG_ORDER_SFC);
    ~Domain();
    Domain(const Domain&) = delete;
    Domain& operator=(const Domain&) = delete;
    void zero_bc_iv();

    hcfd_ctx* ctx() const { return ctx_; }
    // BC ordinal of the boundary colour whose faces start at F1 (-1 if none).
    int bc_of_face_range(int F1) const;
    // Records one BC's type and primitive state and uploads the BC table.
    void register_bc(int bc, int type, const double QBC[NVARS], double gma);

   private:
    hcfd_ctx* ctx_ = nullptr;
    std::vector<int> bctype_;
    std::vector<double> qbc_;
};

// Throws std::runtime_error with hcfd_last_error() when rc != HCFD_OK.
void check(int rc, const char* what);

}  // namespace MI355X
