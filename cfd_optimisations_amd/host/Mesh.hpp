// Mesh.hpp -- .hmsh mesh with the members of the reference's Mesh
// (includes/Mesh.hpp:4-22): raw CSR arrays plus the cell->face CSR.
#pragma once
#include <string>

class Mesh {
   public:
    int dim, Nn, Nc, Nf, Ncol, NBCcol, N_pnames;
    std::string* pnames;
    int *coltag, *coloff;
    double* nodes;
    int *fn_con, *fc_con;
    int *cf_con, *cf_off;
    int *cn_off, *cn_con;
    int *cc_off, *cc_con;

    // Mesh::Mesh (src/Mesh.cpp:7-52); throws std::runtime_error on an unreadable file.
    explicit Mesh(const char* fname);
    ~Mesh();
    Mesh(const Mesh&) = delete;
    Mesh& operator=(const Mesh&) = delete;

    void colour_faces();

   private:
    void compute_cf_con();
};
