// ref_driver.cpp -- C entry points over the reference's own Base classes.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline; never linked into the
// product).  It is compiled together with the reference's sources, where they
// lie, by oracle/Makefile into oracle/_ref/.  Nothing of the reference is
// copied: this file only constructs the reference's Mesh/Base::Domain/
// Base::GaussGreen/Base::Limiter/Base::Flux/Base::BoundaryCondition objects
// and calls their methods.
//
// src/Base/Euler.cpp is not compiled: its constructor builds a Writer, and
// src/Writer.cpp needs <format> (GCC >= 13; this image has GCC 11).  The step
// loop of Euler::run (src/Base/Euler.cpp:85-97), Euler::initialise (:62-77),
// Euler::gather_BCs (:42-60) and the update_BC_* loops (:138-154) are restated
// below in the same order; every arithmetic kernel is the reference's own.
#include <cstring>
#include <vector>

#include "Mesh.hpp"
#include "Base/Domain.hpp"
#include "Base/GradientSchemes.hpp"
#include "Base/Limiters.hpp"
#include "Base/Flux.hpp"
#include "Base/BoundaryConditions.hpp"

namespace {

struct RefCase {
    Mesh* mesh = nullptr;
    Base::Domain* dom = nullptr;
    Base::GaussGreen* grads = nullptr;
    Base::Limiter* lim = nullptr;
    Base::Flux* flux = nullptr;
    std::vector<Base::BoundaryCondition*> bcs;
    int Nc = 0;
    double* U[NVARS];
    double* R[NVARS];
    double* phi[NVARS];
    double* G[NVARS];
};

void to_soa(int Nc, const double* flat, double* const* arr, int width) {
    for (int v = 0; v < NVARS; v++) std::memcpy(arr[v], flat + (size_t)v * width * Nc, sizeof(double) * width * Nc);
}
void from_soa(int Nc, double* flat, double* const* arr, int width) {
    for (int v = 0; v < NVARS; v++) std::memcpy(flat + (size_t)v * width * Nc, arr[v], sizeof(double) * width * Nc);
}

}  // namespace

extern "C" {

// bcmap is indexed by colour tag (like Test.cpp's BCmap, src/Test.cpp:64) and
// qbc by BC ordinal (like QBC, src/Test.cpp:57); gather_BCs, Euler.cpp:42-60.
void* ref_create(const char* hmsh, double gma, const int* bcmap, const double* qbc) {
    RefCase* c = new RefCase();
    c->mesh = new Mesh(hmsh);
    c->dom = new Base::Domain(*c->mesh);
    c->grads = new Base::GaussGreen(*c->dom);
    c->lim = new Base::Limiter(*c->dom);
    c->flux = new Base::Flux(NVARS, gma, *c->dom);
    c->Nc = c->dom->Nc;
    for (int v = 0; v < NVARS; v++) {
        c->U[v] = new double[c->Nc]();
        c->R[v] = new double[c->Nc]();
        c->phi[v] = new double[c->Nc]();
        c->G[v] = new double[NDIMS * c->Nc]();
    }
    int fcolBC = c->dom->Ncol - c->dom->NBC;
    for (int fcol = fcolBC; fcol < c->dom->Ncol; fcol++) {
        int bc = fcol - fcolBC;
        Base::BoundaryCondition* b = nullptr;
        switch ((Base::BCtype)bcmap[c->mesh->coltag[fcol]]) {
            case Base::SYMMETRY: b = new Base::Symmetry(*c->dom); break;
            case Base::WALL: b = new Base::Wall(*c->dom); break;
            case Base::SUPINLET: b = new Base::SupInlet(*c->dom); break;
            case Base::SUBINLET: b = new Base::SubInlet(*c->dom); break;
            case Base::SUPOUTLET: b = new Base::SupOutlet(*c->dom); break;
            case Base::SUBOUTLET: b = new Base::SubOutlet(*c->dom); break;
        }
        b->init(NVARS, c->mesh->coloff[fcol], c->mesh->coloff[fcol + 1], gma, qbc + NVARS * bc);
        c->bcs.push_back(b);
    }
    return c;
}

void ref_destroy(void* h) {
    RefCase* c = (RefCase*)h;
    for (auto* b : c->bcs) delete b;
    for (int v = 0; v < NVARS; v++) {
        delete[] c->U[v]; delete[] c->R[v]; delete[] c->phi[v]; delete[] c->G[v];
    }
    delete c->flux; delete c->lim; delete c->grads; delete c->dom; delete c->mesh;
    delete c;
}

void ref_sizes(void* h, int* out) {
    RefCase* c = (RefCase*)h;
    out[0] = c->dom->Nc; out[1] = c->dom->Nf; out[2] = c->dom->Ncol; out[3] = c->dom->NBC;
}

// Geometry as Base::Domain holds it (rf [Nf][2][2], nf/xf [Nf][2], xc [Nc][2]).
void ref_geometry(void* h, int* C, double* xf, double* xc, double* rf, double* nf, double* Af, double* If, double* iv) {
    RefCase* c = (RefCase*)h;
    size_t Nf = c->dom->Nf, Nc = c->dom->Nc;
    std::memcpy(C, c->dom->C, sizeof(int) * 2 * Nf);
    std::memcpy(xf, c->dom->xf, sizeof(double) * 2 * Nf);
    std::memcpy(xc, c->dom->xc, sizeof(double) * 2 * Nc);
    std::memcpy(rf, c->dom->rf, sizeof(double) * 4 * Nf);
    std::memcpy(nf, c->dom->nf, sizeof(double) * 2 * Nf);
    std::memcpy(Af, c->dom->Af, sizeof(double) * Nf);
    std::memcpy(If, c->dom->If, sizeof(double) * Nf);
    std::memcpy(iv, c->dom->iv, sizeof(double) * Nc);
}

void ref_cf(void* h, int* cf_off, int* cf_con) {
    RefCase* c = (RefCase*)h;
    std::memcpy(cf_off, c->mesh->cf_off, sizeof(int) * (c->mesh->Nc + 1));
    std::memcpy(cf_con, c->mesh->cf_con, sizeof(int) * c->mesh->cf_off[c->mesh->Nc]);
}

void ref_zero_bc_iv(void* h) { ((RefCase*)h)->dom->zero_bc_iv(); }

void ref_gradient(void* h, const double* U, double* G) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, U, c->U, 1);
    c->grads->compute_gradient(NVARS, c->U, c->G);
    from_soa(c->Nc, G, c->G, NDIMS);
}

void ref_limiter(void* h, const double* U, const double* G, double* phi) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, U, c->U, 1);
    to_soa(c->Nc, G, c->G, NDIMS);
    c->lim->limit_gradient(NVARS, c->U, c->G, c->phi);
    from_soa(c->Nc, phi, c->phi, 1);
}

void ref_bc_grads(void* h, double* G) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, G, c->G, NDIMS);
    for (auto* b : c->bcs) b->update_grads(c->G);
    from_soa(c->Nc, G, c->G, NDIMS);
}

void ref_bc_limiter(void* h, double* phi) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, phi, c->phi, 1);
    for (auto* b : c->bcs) b->update_limiter(c->phi);
    from_soa(c->Nc, phi, c->phi, 1);
}

// R accumulates into the caller's array, then R *= iv (Flux.cpp:9-50).
void ref_flux(void* h, const double* U, const double* G, const double* phi, double* R) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, U, c->U, 1);
    to_soa(c->Nc, G, c->G, NDIMS);
    to_soa(c->Nc, phi, c->phi, 1);
    to_soa(c->Nc, R, c->R, 1);
    c->flux->ComputeFlux(c->U, c->G, c->phi, c->R);
    from_soa(c->Nc, R, c->R, 1);
}

void ref_bc_cells(void* h, double* U) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, U, c->U, 1);
    for (auto* b : c->bcs) b->update_cells(c->U);
    from_soa(c->Nc, U, c->U, 1);
}

// Euler::initialise, Euler.cpp:62-77 (uniform state, then BC ghost cells).
void ref_initialise(void* h, const double* QIC, double gma, double* U) {
    RefCase* c = (RefCase*)h;
    double UIC[NVARS];
    UIC[0] = QIC[0];
    for (int v = 1; v < NVARS; v++) UIC[v] = QIC[0] * QIC[v];
    UIC[NDIMS + 1] = QIC[NDIMS + 1] / (gma - 1) + 0.5 * QIC[0] * dot(&QIC[1], &QIC[1]);
    for (int v = 0; v < NVARS; v++)
        for (int i = 0; i < c->Nc; i++) c->U[v][i] = UIC[v];
    for (auto* b : c->bcs) b->update_cells(c->U);
    from_soa(c->Nc, U, c->U, 1);
}

// The step loop of Euler::run, Euler.cpp:85-97, for nsteps steps of fixed dt;
// U ([4][Nc]) in/out.  Requires ref_zero_bc_iv first, as the Euler ctor does.
void ref_run(void* h, double* U, int nsteps, double dt) {
    RefCase* c = (RefCase*)h;
    to_soa(c->Nc, U, c->U, 1);
    for (int v = 0; v < NVARS; v++) std::memset(c->R[v], 0, sizeof(double) * c->Nc);
    for (int s = 0; s < nsteps; s++) {
        c->grads->compute_gradient(NVARS, c->U, c->G);
        c->lim->limit_gradient(NVARS, c->U, c->G, c->phi);
        for (auto* b : c->bcs) b->update_grads(c->G);
        for (auto* b : c->bcs) b->update_limiter(c->phi);
        c->flux->ComputeFlux(c->U, c->G, c->phi, c->R);
        for (int v = 0; v < NVARS; v++)
            for (int i = 0; i < c->Nc; i++) c->U[v][i] += dt * c->R[v][i];
        for (int v = 0; v < NVARS; v++) setarr(0, c->Nc, 0.0, c->R[v]);
        for (auto* b : c->bcs) b->update_cells(c->U);
    }
    from_soa(c->Nc, U, c->U, 1);
}

// Euler::compute_prims, Euler.cpp:115-136, restated (Euler.cpp is not built);
// compiled with the same flags as the reference kernels.
void ref_prims(void* h, double gma, const double* Uflat, double* Qflat) {
    RefCase* c = (RefCase*)h;
    const int Nc = c->Nc;
    const double* U[NVARS];
    double* Q[NVARS];
    for (int v = 0; v < NVARS; v++) { U[v] = Uflat + (size_t)v * Nc; Q[v] = Qflat + (size_t)v * Nc; }
    for (int i = 0; i < Nc; i++) Q[0][i] = U[0][i];
    for (int var = 1; var < NVARS; var++) {
        if (var == NDIMS + 1) continue;
        for (int i = 0; i < Nc; i++) Q[var][i] = U[var][i] / U[0][i];
    }
    double gmo = gma - 1;
    for (int i = 0; i < Nc; i++) {
        double u[NDIMS];
        for (int dim = 0; dim < NDIMS; dim++) u[dim] = Q[dim + 1][i];
        Q[NDIMS + 1][i] = gmo * (U[NDIMS + 1][i] - 0.5 * Q[0][i] * dot(u, u));
    }
}

}  // extern "C"
