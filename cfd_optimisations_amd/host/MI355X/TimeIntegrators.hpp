// MI355X::Euler -- the shape of includes/Base/TimeIntegrators.hpp:11-45.  The
// state lives on the GPU; run() drives the fused device step between outputs.
#pragma once
#include <string>

#include "Global.hpp"
#include "MI355X/BoundaryConditions.hpp"
#include "MI355X/Domain.hpp"
#include "MI355X/Flux.hpp"
#include "MI355X/GradientSchemes.hpp"
#include "MI355X/Limiters.hpp"
#include "Writer.hpp"

namespace MI355X {

class Euler {
   public:
    Euler(int nvars, Domain& domain, const Flux& flux, const GradientScheme& grads, Limiter& limiter);
    ~Euler();
    Euler(const Euler&) = delete;
    Euler& operator=(const Euler&) = delete;
    void gather_BCs(const BCtype* BCmap, const double QBC[][NVARS]);
    void initialise(const double QIC[NVARS]);

    void run(double T, double dt, double dt_out);
    void compute_prims();

    // Per-cell primitive initial state (Euler::initialise's arithmetic cell by
    // cell, then the BC ghosts): Q[v][i], the mesh's cell order.
    void initialise_prims(const double* const Q[NVARS]);
    // run() with the CFL time step (no reference counterpart: the reference steps
    // a fixed dt): each step's dt is cfl times the CFL bound of its state; files
    // are written as run() writes them, when the clock passes a multiple of dt_out.
    void run_cfl(double T, double cfl, double dt_out);

    // Output file stem (the reference writes ../sols/<Layout>, src/Base/Euler.cpp:22).
    void set_output(const std::string& fname);
    // Stages per step of run(): 1 forward Euler (the reference's), 2 SSP-RK2,
    // 3 SSP-RK3 (no reference counterpart; hcfd_set_integrator).
    void set_stages(int stages);
    // Start from a .hsln file written by run() (primitives of all cells) instead
    // of initialise(); the file must belong to this mesh.
    void restart(const std::string& hsln_path);
    // Checkpoint / resume, no reference counterpart: the conserved state, the
    // clock and the last dt, bit for bit (hcfd_checkpoint_save / _load); a
    // later run() continues from the loaded state.
    void checkpoint(const std::string& path);
    void resume(const std::string& path);
    // Conserved state copied back from the GPU into U.
    const double* const* state();

   private:
    const int nvars, Nc, Nf;
    Domain& domain;
    const Flux& flux;
    const GradientScheme& grads;
    Limiter& limiter;

    const int NBC;
    BoundaryCondition** BCs;

    double* Q[NVARS];
    double* U[NVARS];
    double* R[NVARS];
    double* phi[NVARS];
    double* G[NVARS];

    Writer* writer;
};

}  // namespace MI355X
