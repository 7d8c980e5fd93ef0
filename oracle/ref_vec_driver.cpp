// ref_vec_driver.cpp -- C entry points over the reference's VEC layout, its
// fastest valid CPU layout (SIMD face chunks of VECLEN; built for AVX2, where
// VECLEN = 4: an AVX-512 build scatters conflicting lanes and ends in NaN).
//
// TEST INFRASTRUCTURE ONLY (CPU baseline and checker; never linked into the
// product).  Compiled together with the reference's own sources, where they
// lie, by oracle/Makefile into oracle/_ref/.  Nothing of the reference is
// copied: this file constructs Mesh / VEC::Domain / VEC::GaussGreen /
// VEC::Limiter / VEC::Flux / VEC::BoundaryCondition objects and calls their
// methods.  src/VEC/Euler.cpp is not compiled (its constructor builds a Writer,
// which needs <format>); its arrays (Nc + 1 entries, the sentinel cell of the
// padded face chunks, src/VEC/Euler.cpp:10-16) and its step loop
// (src/VEC/Euler.cpp:80-104) are restated in the same order.
#include <immintrin.h>

#include <cstring>
#include <vector>

#include "Mesh.hpp"
#include "VEC/BoundaryConditions.hpp"
#include "VEC/Domain.hpp"
#include "VEC/Flux.hpp"
#include "VEC/GradientSchemes.hpp"
#include "VEC/Limiters.hpp"

namespace {

struct VecCase {
    Mesh* mesh = nullptr;
    VEC::Domain* dom = nullptr;
    VEC::GaussGreen* grads = nullptr;
    VEC::Limiter* lim = nullptr;
    VEC::Flux* flux = nullptr;
    std::vector<VEC::BoundaryCondition*> bcs;
    int Nc = 0;
    double* U[NVARS];
    double* R[NVARS];
    double* phi[NVARS];
    double* G[NVARS];
};

double* alloc(size_t n) {
    double* p = (double*)_mm_malloc(sizeof(double) * n, ALIGN);
    std::memset(p, 0, sizeof(double) * n);
    return p;
}

}  // namespace

extern "C" {

// bcmap by colour tag, qbc by BC ordinal (src/VEC/Euler.cpp:43-61).
void* vref_create(const char* hmsh, double gma, const int* bcmap, const double* qbc) {
    VecCase* c = new VecCase();
    c->mesh = new Mesh(hmsh);
    c->dom = new VEC::Domain(*c->mesh);
    c->grads = new VEC::GaussGreen(*c->dom);
    c->lim = new VEC::Limiter(*c->dom);
    c->flux = new VEC::Flux(NVARS, gma, *c->dom);
    c->Nc = c->dom->Nc;
    for (int v = 0; v < NVARS; v++) {
        c->U[v] = alloc(c->Nc + 1);
        c->R[v] = alloc(c->Nc + 1);
        c->phi[v] = alloc(c->Nc + 1);
        c->G[v] = alloc((size_t)NDIMS * c->Nc + NDIMS);
    }
    const int fcolBC = c->dom->Ncol - c->dom->NBC;
    for (int fcol = fcolBC; fcol < c->dom->Ncol; fcol++) {
        const int bc = fcol - fcolBC;
        VEC::BoundaryCondition* b = nullptr;
        switch ((VEC::BCtype)bcmap[c->mesh->coltag[fcol]]) {
            case VEC::SYMMETRY: b = new VEC::Symmetry(*c->dom); break;
            case VEC::WALL: b = new VEC::Wall(*c->dom); break;
            case VEC::SUPINLET: b = new VEC::SupInlet(*c->dom); break;
            case VEC::SUBINLET: b = new VEC::SubInlet(*c->dom); break;
            case VEC::SUPOUTLET: b = new VEC::SupOutlet(*c->dom); break;
            case VEC::SUBOUTLET: b = new VEC::SubOutlet(*c->dom); break;
        }
        b->init(NVARS, c->mesh->coloff[fcol], c->mesh->coloff[fcol + 1], gma, qbc + NVARS * bc);
        c->bcs.push_back(b);
    }
    c->dom->zero_bc_iv();  // as the Euler constructor does
    return c;
}

void vref_destroy(void* h) {
    VecCase* c = (VecCase*)h;
    for (auto* b : c->bcs) delete b;
    for (int v = 0; v < NVARS; v++) {
        _mm_free(c->U[v]); _mm_free(c->R[v]); _mm_free(c->phi[v]); _mm_free(c->G[v]);
    }
    delete c->flux; delete c->lim; delete c->grads; delete c->dom; delete c->mesh;
    delete c;
}

int vref_veclen(void) { return VECLEN; }

// The step loop of src/VEC/Euler.cpp:85-97 for nsteps steps of fixed dt; U is
// [4][Nc] in and out (the sentinel entries stay zero).
void vref_run(void* h, double* U, int nsteps, double dt) {
    VecCase* c = (VecCase*)h;
    const int Nc = c->Nc;
    for (int v = 0; v < NVARS; v++) {
        std::memcpy(c->U[v], U + (size_t)v * Nc, sizeof(double) * Nc);
        std::memset(c->R[v], 0, sizeof(double) * (Nc + 1));
    }
    for (int s = 0; s < nsteps; s++) {
        c->grads->compute_gradient(NVARS, c->U, c->G);
        c->lim->limit_gradient(NVARS, c->U, c->G, c->phi);
        for (auto* b : c->bcs) b->update_grads(c->G);
        for (auto* b : c->bcs) b->update_limiter(c->phi);
        c->flux->ComputeFlux(c->U, c->G, c->phi, c->R);
        for (int v = 0; v < NVARS; v++)
            for (int i = 0; i < Nc; i++) c->U[v][i] += dt * c->R[v][i];
        for (int v = 0; v < NVARS; v++) setarr(0, Nc, 0.0, c->R[v]);
        for (auto* b : c->bcs) b->update_cells(c->U);
    }
    for (int v = 0; v < NVARS; v++) std::memcpy(U + (size_t)v * Nc, c->U[v], sizeof(double) * Nc);
}

// Euler::compute_prims, src/VEC/Euler.cpp:114-135, restated (Euler.cpp is not
// built); compiled with the reference kernels' flags, as the shipped files were.
void vref_prims(void* h, double gma, const double* Uflat, double* Qflat) {
    VecCase* c = (VecCase*)h;
    const int Nc = c->Nc;
    const double* U[NVARS];
    double* Q[NVARS];
    for (int v = 0; v < NVARS; v++) { U[v] = Uflat + (size_t)v * Nc; Q[v] = Qflat + (size_t)v * Nc; }
    for (int i = 0; i < Nc; i++) Q[0][i] = U[0][i];
    for (int var = 1; var < NVARS; var++) {
        if (var == NDIMS + 1) continue;
        for (int i = 0; i < Nc; i++) Q[var][i] = U[var][i] / U[0][i];
    }
    const double gmo = gma - 1;
    for (int i = 0; i < Nc; i++) {
        double u[NDIMS];
        for (int dim = 0; dim < NDIMS; dim++) u[dim] = Q[dim + 1][i];
        Q[NDIMS + 1][i] = gmo * (U[NDIMS + 1][i] - 0.5 * Q[0][i] * dot(u, u));
    }
}

}  // extern "C"
