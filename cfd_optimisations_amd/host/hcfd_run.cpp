// hcfd_run.cpp -- the reference's Test.cpp case (src/Test.cpp:50-74) driven
// through the MI355X drop-in classes:
//
//   hcfd_run [mesh.hmsh] [output stem] [steps] [--stages] [--outputs K] [--rk S]
//            [--restart file.hsln] [--mixed]
//
// Default: ../wedge.hmsh, ../sols/MI355X, 1000 steps of dt = 1e-3, writing
// <stem>_0.hsln and <stem>_1.hsln as the reference writes sols/Base_{0,1}.hsln.
// --stages runs the steps through the per-stage classes on host arrays
// (GaussGreen, Limiter, BC copies, Flux, update, BC cells) instead of the
// fused device step.  --outputs K writes K files after the initial one
// (dt_out = T / K); --rk S steps with S-stage SSP-RK; --restart starts from a
// .hsln file instead of the uniform initial state; --mixed builds the Domain
// for the FP32-mixed path (MI355X::Domain(mesh, HCFD_MIXED); the per-stage
// classes of --stages are FP64 only).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <vector>

#include "MI355X/BoundaryConditions.hpp"
#include "MI355X/Domain.hpp"
#include "MI355X/Flux.hpp"
#include "MI355X/GradientSchemes.hpp"
#include "MI355X/Limiters.hpp"
#include "MI355X/TimeIntegrators.hpp"
#include "Mesh.hpp"

int main(int argc, char** argv) {
    try {
        const char* mesh_path = argc > 1 ? argv[1] : "../wedge.hmsh";
        const std::string out = argc > 2 ? argv[2] : "../sols/MI355X";
        const int niter = argc > 3 ? std::atoi(argv[3]) : 1000;
        bool stages = false, mixed = false;
        int nout = 1, rk = 1;
        std::string restart;
        for (int k = 4; k < argc; k++) {
            if (std::strcmp(argv[k], "--stages") == 0) stages = true;
            else if (std::strcmp(argv[k], "--outputs") == 0 && k + 1 < argc) nout = std::atoi(argv[++k]);
            else if (std::strcmp(argv[k], "--rk") == 0 && k + 1 < argc) rk = std::atoi(argv[++k]);
            else if (std::strcmp(argv[k], "--restart") == 0 && k + 1 < argc) restart = argv[++k];
            else if (std::strcmp(argv[k], "--mixed") == 0) mixed = true;
            else throw std::runtime_error(std::string("unknown option ") + argv[k]);
        }
        if (nout < 1) throw std::runtime_error("--outputs must be at least 1");
        if (mixed && stages) throw std::runtime_error("--stages runs the FP64 per-stage classes; drop --mixed");

        // the case of src/Test.cpp:51-64
        const int nvars = 4;
        const double gma = 1.4;
        const double M = 2.0;
        const double p = 1.0 / (gma * M * M);
        const double QIC[] = {1, 1, 0, p};
        const double QBC[][NVARS] = {{1, 1, 0, p}, {1, 1, 0, p}, {1, 1, 0, p}, {1, 1, 0, p}};
        const double dt = 1e-3;
        const MI355X::BCtype BCmap[] = {MI355X::SYMMETRY, MI355X::SYMMETRY, MI355X::SYMMETRY, MI355X::SUPINLET,
                                        MI355X::SUPOUTLET};

        Mesh mesh(mesh_path);
        MI355X::Domain domain(mesh, mixed ? HCFD_MIXED : HCFD_FP64);
        MI355X::GaussGreen grads(domain);
        MI355X::Limiter limiter(domain);
        MI355X::Flux flux(nvars, gma, domain);
        MI355X::Euler intg(nvars, domain, flux, grads, limiter);
        intg.set_output(out);
        intg.gather_BCs(BCmap, QBC);
        intg.initialise(QIC);
        if (!restart.empty()) intg.restart(restart);
        intg.set_stages(rk);

        auto t1 = std::chrono::steady_clock::now();
        if (!stages) {
            intg.run(niter * dt, dt, niter * dt / nout);
        } else {
            // Euler::run's loop body (src/Base/Euler.cpp:85-97) stage by stage
            const int Nc = domain.Nc;
            std::vector<double> Ub(4 * (size_t)Nc), Gb(8 * (size_t)Nc), Pb(4 * (size_t)Nc), Rb(4 * (size_t)Nc, 0.0);
            double *U[4], *G[4], *phi[4], *R[4];
            for (int v = 0; v < 4; v++) {
                U[v] = &Ub[(size_t)v * Nc];
                G[v] = &Gb[(size_t)v * 2 * Nc];
                phi[v] = &Pb[(size_t)v * Nc];
                R[v] = &Rb[(size_t)v * Nc];
            }
            const double* const* U0 = intg.state();
            for (int v = 0; v < 4; v++) std::memcpy(U[v], U0[v], sizeof(double) * Nc);
            intg.compute_prims();
            // one BoundaryCondition object per colour, as gather_BCs makes them
            std::vector<MI355X::BoundaryCondition*> bcs;
            for (int fcol = domain.Ncol - domain.NBC; fcol < domain.Ncol; fcol++) {
                const int b = fcol - (domain.Ncol - domain.NBC);
                MI355X::BoundaryCondition* bc = nullptr;
                switch (BCmap[mesh.coltag[fcol]]) {
                    case MI355X::SYMMETRY: bc = new MI355X::Symmetry(domain); break;
                    case MI355X::SUPINLET: bc = new MI355X::SupInlet(domain); break;
                    default: bc = new MI355X::SupOutlet(domain); break;
                }
                bc->init(nvars, mesh.coloff[fcol], mesh.coloff[fcol + 1], gma, QBC[b]);
                bcs.push_back(bc);
            }
            double t = 0.0;
            while (t < niter * dt) {
                grads.compute_gradient(nvars, U, G);
                limiter.limit_gradient(nvars, U, G, phi);
                for (auto* b : bcs) b->update_grads(G);
                for (auto* b : bcs) b->update_limiter(phi);
                flux.ComputeFlux(U, G, phi, R);
                for (size_t k = 0; k < Ub.size(); k++) Ub[k] += dt * Rb[k];
                std::fill(Rb.begin(), Rb.end(), 0.0);
                for (auto* b : bcs) b->update_cells(U);
                t += dt;
            }
            for (auto* b : bcs) delete b;
            const double* const Uc[4] = {U[0], U[1], U[2], U[3]};
            MI355X::check(hcfd_set_state(domain.ctx(), Uc), "hcfd_set_state");
            intg.compute_prims();
            // write the final state like run() would
            std::vector<double> Qb(4 * (size_t)Nc);
            double* Q[4];
            for (int v = 0; v < 4; v++) Q[v] = &Qb[(size_t)v * Nc];
            MI355X::check(hcfd_get_prims(domain.ctx(), Q), "hcfd_get_prims");
            const double* const Qc[4] = {Q[0], Q[1], Q[2], Q[3]};
            MI355X::check(hcfd_write_hsln((out + "_1.hsln").c_str(), Nc, Qc, nullptr), "hcfd_write_hsln");
        }
        auto t2 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t2 - t1).count();
        std::printf("MI355X%s: %d steps on %d cells in %.4f s (%.3e cell-updates/s incl. ghosts)\n",
                    stages ? " (per-stage classes)" : "", niter, domain.Nc, s, (double)domain.Nc * niter / s);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "hcfd_run: %s\n", e.what());
        return 1;
    }
}
