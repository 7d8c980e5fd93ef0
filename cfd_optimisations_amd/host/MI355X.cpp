// MI355X.cpp -- the drop-in classes of namespace MI355X over the C ABI.
//
// Per-stage methods (GaussGreen::compute_gradient, Limiter::limit_gradient,
// Flux::ComputeFlux, BoundaryCondition::update_*) take and return host
// arrays exactly like Base's, moving them to and from the device around one
// kernel; Euler keeps the state on the device and runs the fused step.
#include <cstring>
#include <stdexcept>
#include <string>

#include "MI355X/BoundaryConditions.hpp"
#include "MI355X/Domain.hpp"
#include "MI355X/Flux.hpp"
#include "MI355X/GradientSchemes.hpp"
#include "MI355X/Limiters.hpp"
#include "MI355X/TimeIntegrators.hpp"

namespace MI355X {

void check(int rc, const char* what) {
    if (rc != HCFD_OK) throw std::runtime_error(std::string(what) + ": " + hcfd_last_error());
}

// ---- Domain (src/Base/Domain.cpp:4-154) -------------------------------------

Domain::Domain(Mesh& mesh, int precision, int flags)
    : Nf(mesh.Nf), Nc(mesh.Nc), Ncol(mesh.Ncol), NBC(mesh.NBCcol), mesh(mesh) {
    hcfd_mesh_view v{mesh.dim, mesh.Nn, mesh.Ncol, mesh.NBCcol, mesh.Nc, mesh.Nf, mesh.coltag, mesh.coloff,
                     mesh.nodes, mesh.fn_con, mesh.fc_con, mesh.cn_off, mesh.cn_con};
    check(hcfd_create(&v, precision, flags, &ctx_), "hcfd_create");
    C = new int[Nf][2];
    rf = new double[Nf][2][NDIMS];
    xf = new double[Nf][NDIMS];
    nf = new double[Nf][NDIMS];
    Af = new double[Nf];
    If = new double[Nf];
    xc = new double[Nc][NDIMS];
    iv = new double[Nc];
    std::memcpy(C, mesh.fc_con, sizeof(int) * 2 * Nf);
    check(hcfd_get_geometry(ctx_, &xf[0][0], &xc[0][0], &rf[0][0][0], &nf[0][0], Af, If, iv), "hcfd_get_geometry");
    // the device applies zero_bc_iv at creation; the host copy keeps the
    // pre-zero values (ghost 1/V = its interior neighbour's) until asked
    const int F_BC = mesh.coloff[Ncol - NBC];
    for (int f = F_BC; f < Nf; f++) iv[C[f][1]] = iv[C[f][0]];
    bctype_.assign(NBC, SUPOUTLET);
    qbc_.assign((size_t)NVARS * NBC, 1.0);
    check(hcfd_set_gamma(ctx_, 1.4), "hcfd_set_gamma");
}

Domain::~Domain() {
    hcfd_destroy(ctx_);
    delete[] C;
    delete[] rf;
    delete[] xf;
    delete[] nf;
    delete[] Af;
    delete[] If;
    delete[] xc;
    delete[] iv;
}

void Domain::zero_bc_iv() {
    const int F_BC = mesh.coloff[Ncol - NBC];
    for (int f = F_BC; f < Nf; f++) iv[C[f][1]] = 0.0;
}

int Domain::bc_of_face_range(int F1) const {
    for (int col = Ncol - NBC; col < Ncol; col++)
        if (mesh.coloff[col] == F1) return col - (Ncol - NBC);
    return -1;
}

void Domain::register_bc(int bc, int type, const double QBC[NVARS], double gma) {
    if (bc < 0 || bc >= NBC) throw std::runtime_error("register_bc: no boundary colour starts at this face range");
    bctype_[bc] = type;
    for (int v = 0; v < NVARS; v++) qbc_[(size_t)NVARS * bc + v] = QBC[v];
    check(hcfd_set_gamma(ctx_, gma), "hcfd_set_gamma");
    check(hcfd_set_bcs(ctx_, NBC, bctype_.data(), qbc_.data(), gma), "hcfd_set_bcs");
}

namespace {
void put(hcfd_ctx* c, int which, const double* const X[NVARS]) {
    check(hcfd_stage_set(c, which, X), "hcfd_stage_set");
}
void get(hcfd_ctx* c, int which, double* const X[NVARS]) { check(hcfd_stage_get(c, which, X), "hcfd_stage_get"); }
}  // namespace

// ---- gradients (src/Base/GaussGreen.cpp:4-40) ---------------------------------

GradientScheme::GradientScheme(const Domain& domain) : Nc(domain.Nc), Nf(domain.Nf), domain(domain) {}

void GradientScheme::compute_gradient(int, const double* const[NVARS], double*[NVARS]) const {}

void GaussGreen::compute_gradient(int nvars, const double* const u[NVARS], double* grad_u[NVARS]) const {
    if (nvars != NVARS) throw std::runtime_error("GaussGreen: nvars must be 4");
    put(domain.ctx(), HCFD_ARR_U, u);
    check(hcfd_stage_gradient(domain.ctx()), "hcfd_stage_gradient");
    get(domain.ctx(), HCFD_ARR_G, grad_u);
}

// ---- limiter (src/Base/Limiter.cpp:19-77) -------------------------------------

Limiter::Limiter(const Domain& domain) : domain(domain) {}

void Limiter::limit_gradient(const int nvars, const double* const U[NVARS], const double* const G[NVARS],
                             double* phi[NVARS]) {
    if (nvars != NVARS) throw std::runtime_error("Limiter: nvars must be 4");
    put(domain.ctx(), HCFD_ARR_U, U);
    put(domain.ctx(), HCFD_ARR_G, G);
    check(hcfd_stage_limiter(domain.ctx()), "hcfd_stage_limiter");
    get(domain.ctx(), HCFD_ARR_PHI, phi);
}

// ---- flux (src/Base/Flux.cpp:9-109) ---------------------------------------------

Flux::Flux(int nvars, double gma, const Domain& domain) : nvars(nvars), gma(gma), domain(domain) {
    check(hcfd_set_gamma(domain.ctx(), gma), "hcfd_set_gamma");
}

void Flux::ComputeFlux(const double* const U[NVARS], const double* const G[NVARS], const double* const phi[NVARS],
                       double* R[NVARS]) const {
    hcfd_ctx* c = domain.ctx();
    put(c, HCFD_ARR_U, U);
    put(c, HCFD_ARR_G, G);
    put(c, HCFD_ARR_PHI, phi);
    put(c, HCFD_ARR_R, R);
    check(hcfd_stage_flux(c), "hcfd_stage_flux");
    get(c, HCFD_ARR_R, R);
}

// ---- boundary conditions (src/Base/BoundaryConditions/*.cpp) -------------------

BoundaryCondition::BoundaryCondition(const Domain& domain) : nvars(NVARS), Nc(domain.Nc), F1(0), F2(0), gma(1.4),
                                                             domain(domain) {}

void BoundaryCondition::init(int nvars, int F1, int F2, double gma, const double QBC[NVARS]) {
    this->nvars = nvars;
    this->F1 = F1;
    this->F2 = F2;
    this->gma = gma;
    for (int v = 0; v < nvars; v++) this->QBC[v] = QBC[v];
    // BoundaryCondition::init's UBC (src/Base/BoundaryConditions/Base.cpp:18-22)
    UBC[0] = QBC[0];
    for (int v = 1; v < nvars; v++) UBC[v] = QBC[0] * QBC[v];
    UBC[NDIMS + 1] = QBC[NDIMS + 1] / (gma - 1) + 0.5 * QBC[0] * dot(&QBC[1], &QBC[1]);
    bc = domain.bc_of_face_range(F1);
    const_cast<Domain&>(domain).register_bc(bc, type(), QBC, gma);
}

void BoundaryCondition::update_cells(double* U[NVARS]) {
    put(domain.ctx(), HCFD_ARR_U, U);
    check(hcfd_stage_bc(domain.ctx(), 0, bc), "hcfd_stage_bc");
    get(domain.ctx(), HCFD_ARR_U, U);
}

void BoundaryCondition::update_grads(double* G[NVARS]) {
    put(domain.ctx(), HCFD_ARR_G, G);
    check(hcfd_stage_bc(domain.ctx(), 1, bc), "hcfd_stage_bc");
    get(domain.ctx(), HCFD_ARR_G, G);
}

void BoundaryCondition::update_limiter(double* phi[NVARS]) {
    put(domain.ctx(), HCFD_ARR_PHI, phi);
    check(hcfd_stage_bc(domain.ctx(), 2, bc), "hcfd_stage_bc");
    get(domain.ctx(), HCFD_ARR_PHI, phi);
}

// ---- Euler (src/Base/Euler.cpp:4-154) ----------------------------------------------

Euler::Euler(int nvars, Domain& domain, const Flux& flux, const GradientScheme& grads, Limiter& limiter)
    : nvars(nvars), Nc(domain.Nc), Nf(domain.Nf), domain(domain), flux(flux), grads(grads), limiter(limiter),
      NBC(domain.NBC) {
    BCs = new BoundaryCondition*[NBC]();
    for (int v = 0; v < nvars; v++) {
        Q[v] = new double[Nc]();
        U[v] = new double[Nc]();
        R[v] = new double[Nc]();
        phi[v] = new double[Nc]();
        G[v] = new double[NDIMS * Nc]();
    }
    const std::string varnames[] = {"Density", "Velocity_X", "Velocity_Y", "Pressure"};
    writer = new Writer(nvars, "../sols/MI355X", varnames, Nc, Q);
    domain.zero_bc_iv();
}

Euler::~Euler() {
    for (int bc = 0; bc < NBC; bc++) delete BCs[bc];
    delete[] BCs;
    for (int v = 0; v < nvars; v++) {
        delete[] Q[v];
        delete[] U[v];
        delete[] R[v];
        delete[] phi[v];
        delete[] G[v];
    }
    delete writer;
}

void Euler::set_output(const std::string& fname) {
    const std::string varnames[] = {"Density", "Velocity_X", "Velocity_Y", "Pressure"};
    delete writer;
    writer = new Writer(nvars, fname, varnames, Nc, Q);
}

// Euler::gather_BCs (src/Base/Euler.cpp:42-60): one BC object per boundary colour.
void Euler::gather_BCs(const BCtype* BCmap, const double QBC[][NVARS]) {
    const int fcolBC = domain.Ncol - NBC;
    for (int fcol = fcolBC; fcol < domain.Ncol; fcol++) {
        const int F1 = domain.mesh.coloff[fcol], F2 = domain.mesh.coloff[fcol + 1];
        const int bc = fcol - fcolBC;
        switch (BCmap[domain.mesh.coltag[fcol]]) {
            case SYMMETRY: BCs[bc] = new Symmetry(domain); break;
            case WALL: BCs[bc] = new Wall(domain); break;
            case SUPINLET: BCs[bc] = new SupInlet(domain); break;
            case SUBINLET: BCs[bc] = new SubInlet(domain); break;
            case SUPOUTLET: BCs[bc] = new SupOutlet(domain); break;
            case SUBOUTLET: BCs[bc] = new SubOutlet(domain); break;
            default: throw std::runtime_error("gather_BCs: unknown BC type");
        }
        BCs[bc]->init(nvars, F1, F2, flux.gma, QBC[bc]);
    }
}

// Euler::initialise (src/Base/Euler.cpp:62-77), on the device.
void Euler::set_stages(int stages) { check(hcfd_set_integrator(domain.ctx(), stages), "hcfd_set_integrator"); }

void Euler::restart(const std::string& hsln_path) {
    int n = 0;
    check(hcfd_read_hsln(hsln_path.c_str(), &n, nullptr), "hcfd_read_hsln");
    if (n != Nc)
        throw std::runtime_error("restart: " + hsln_path + " holds " + std::to_string(n) + " cells, the mesh " +
                                 std::to_string(Nc));
    check(hcfd_read_hsln(hsln_path.c_str(), &n, Q), "hcfd_read_hsln");
    const double* const Qc[NVARS] = {Q[0], Q[1], Q[2], Q[3]};
    check(hcfd_initialise_prims(domain.ctx(), Qc), "hcfd_initialise_prims");
    check(hcfd_get_state(domain.ctx(), U), "hcfd_get_state");
}

void Euler::checkpoint(const std::string& path) {
    check(hcfd_checkpoint_save(domain.ctx(), path.c_str()), "hcfd_checkpoint_save");
}

void Euler::resume(const std::string& path) {
    check(hcfd_checkpoint_load(domain.ctx(), path.c_str()), "hcfd_checkpoint_load");
    check(hcfd_get_state(domain.ctx(), U), "hcfd_get_state");
}

void Euler::initialise(const double QIC[NVARS]) {
    check(hcfd_initialise(domain.ctx(), QIC), "hcfd_initialise");
    check(hcfd_get_state(domain.ctx(), U), "hcfd_get_state");
}

// Euler::run (src/Base/Euler.cpp:79-105): the same t += dt bookkeeping and
// output cadence; the steps between outputs are one device call.  An output
// snapshot is taken on the device and copied to host memory on a second stream
// while the next steps run; the file of output k is written by the host after
// output k+1's steps are enqueued, so writing overlaps the device's work.
void Euler::run(double T, double dt, double dt_out) {
    hcfd_ctx* ctx = domain.ctx();
    double t = 0.0;
    double t_out = dt_out;
    int slot = 0;
    bool pending = false;  // slot ^ 1 holds a snapshot not yet written
    auto write = [&](int s) {
        check(hcfd_output_wait(ctx, s, Q, nullptr), "hcfd_output_wait");
        writer->write_soln();
    };
    check(hcfd_output_begin(ctx, slot), "hcfd_output_begin");
    write(slot);
    while (t < T) {
        int n = 0;
        bool out = false;
        do {
            t += dt;
            n++;
            out = (t >= t_out);
        } while (!out && t < T);
        check(hcfd_step_async(ctx, n, dt, 0.0), "hcfd_step_async");
        if (out) {
            slot ^= 1;
            check(hcfd_output_begin(ctx, slot), "hcfd_output_begin");
            if (pending) write(slot ^ 1);
            pending = true;
            t_out += dt_out;
        }
    }
    if (pending) write(slot);
    check(hcfd_sync(ctx, nullptr), "hcfd_sync");
}

void Euler::initialise_prims(const double* const Qin[NVARS]) {
    check(hcfd_initialise_prims(domain.ctx(), Qin), "hcfd_initialise_prims");
    check(hcfd_get_state(domain.ctx(), U), "hcfd_get_state");
}

void Euler::run_cfl(double T, double cfl, double dt_out) {
    hcfd_ctx* ctx = domain.ctx();
    double t = 0.0;
    double t_out = dt_out;
    compute_prims();
    writer->write_soln();
    while (t < T) {
        check(hcfd_step(ctx, 1, 0.0, cfl, &t), "hcfd_step");
        if (t >= t_out) {
            compute_prims();
            writer->write_soln();
            t_out += dt_out;
        }
    }
}

// Euler::compute_prims (src/Base/Euler.cpp:115-136), computed on the device.
void Euler::compute_prims() { check(hcfd_get_prims(domain.ctx(), Q), "hcfd_get_prims"); }

const double* const* Euler::state() {
    check(hcfd_get_state(domain.ctx(), U), "hcfd_get_state");
    return U;
}

}  // namespace MI355X
