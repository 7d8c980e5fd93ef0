// mesh_struct.hpp -- host mesh owned by the library (opaque hcfd_mesh of the C ABI):
// the arrays of a .hmsh file (includes/Mesh.hpp:4-22 of the reference).
#pragma once
#include <string>
#include <vector>

struct hcfd_mesh {
    int dim = 2, Nn = 0, Ncol = 0, NBCcol = 0, Nc = 0, Nf = 0;
    std::vector<std::string> pnames;
    std::vector<int> coltag, coloff;
    std::vector<double> nodes;
    std::vector<int> fn_con, fc_con, cn_off, cn_con, cc_off, cc_con, c_ctypes;
};
