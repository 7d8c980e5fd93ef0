// Mesh.cpp -- reads a .hmsh through the library's checked reader and keeps
// the arrays under the reference Mesh's member names (src/Mesh.cpp:7-117).
#include "Mesh.hpp"

#include <cstring>
#include <stdexcept>
#include <string>

#include "hcfd.h"

namespace {
template <typename T>
T* copy_of(const T* src, size_t n) {
    T* p = new T[n ? n : 1];
    if (n) std::memcpy(p, src, sizeof(T) * n);
    return p;
}
}  // namespace

Mesh::Mesh(const char* fname) {
    hcfd_mesh* m = nullptr;
    if (hcfd_mesh_load(fname, &m) != HCFD_OK) throw std::runtime_error(hcfd_last_error());
    hcfd_mesh_view v;
    hcfd_mesh_view_of(m, &v);
    dim = v.dim;
    Nn = v.Nn;
    Nc = v.Nc;
    Nf = v.Nf;
    Ncol = v.Ncol;
    NBCcol = v.NBCcol;
    N_pnames = 0;
    while (hcfd_mesh_pname(m, N_pnames)) N_pnames++;
    pnames = new std::string[N_pnames ? N_pnames : 1];
    for (int k = 0; k < N_pnames; k++) pnames[k] = hcfd_mesh_pname(m, k);
    coltag = copy_of(v.coltag, Ncol);
    coloff = copy_of(v.coloff, Ncol + 1);
    nodes = copy_of(v.nodes, 2 * (size_t)Nn);
    fn_con = copy_of(v.fn_con, 2 * (size_t)Nf);
    fc_con = copy_of(v.fc_con, 2 * (size_t)Nf);
    cn_off = copy_of(v.cn_off, Nc + 1);
    cn_con = copy_of(v.cn_con, v.cn_off[Nc]);
    hcfd_mesh_free(m);
    // cell -> cell CSR of face neighbours
    cc_off = new int[Nc + 1]();
    for (int f = 0; f < Nf; f++) {
        cc_off[fc_con[2 * f] + 1]++;
        cc_off[fc_con[2 * f + 1] + 1]++;
    }
    for (int i = 0; i < Nc; i++) cc_off[i + 1] += cc_off[i];
    cc_con = new int[cc_off[Nc] ? cc_off[Nc] : 1];
    int* cur = new int[Nc]();
    for (int f = 0; f < Nf; f++) {
        int a = fc_con[2 * f], b = fc_con[2 * f + 1];
        cc_con[cc_off[a] + cur[a]++] = b;
        cc_con[cc_off[b] + cur[b]++] = a;
    }
    delete[] cur;
    compute_cf_con();
}

Mesh::~Mesh() {
    delete[] pnames;
    delete[] coltag;
    delete[] coloff;
    delete[] nodes;
    delete[] cn_con;
    delete[] cn_off;
    delete[] fn_con;
    delete[] fc_con;
    delete[] cc_off;
    delete[] cc_con;
    delete[] cf_off;
    delete[] cf_con;
}

// Cell -> face CSR with each cell's faces in ascending face id (the order
// Mesh::compute_cf_con fills, src/Mesh.cpp:88-117).
void Mesh::compute_cf_con() {
    cf_off = new int[Nc + 1]();
    for (int f = 0; f < Nf; f++) {
        cf_off[fc_con[2 * f] + 1]++;
        cf_off[fc_con[2 * f + 1] + 1]++;
    }
    for (int i = 0; i < Nc; i++) cf_off[i + 1] += cf_off[i];
    cf_con = new int[cf_off[Nc] ? cf_off[Nc] : 1];
    int* n = new int[Nc]();
    for (int f = 0; f < Nf; f++) {
        int l = fc_con[2 * f], r = fc_con[2 * f + 1];
        cf_con[cf_off[l] + n[l]++] = f;
        cf_con[cf_off[r] + n[r]++] = f;
    }
    delete[] n;
}

// The working face colouring is built on the device by MI355X::Domain
// (hcfd_get_face_colours); the host Mesh keeps the reference's BC-tag grouping.
void Mesh::colour_faces() {}
