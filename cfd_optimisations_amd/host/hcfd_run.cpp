// hcfd_run.cpp -- the reference's Test.cpp case (src/Test.cpp:50-74) driven
// through the MI355X drop-in classes:
//
//   hcfd_run [mesh.hmsh] [output stem] [steps] [--stages] [--outputs K] [--rk S]
//            [--restart file.hsln] [--resume file.ckp] [--checkpoint file.ckp] [--mixed]
//   hcfd_run --case case.ini [--dry-run]
//
// Default: ../wedge.hmsh, ../sols/MI355X, 1000 steps of dt = 1e-3, writing
// <stem>_0.hsln and <stem>_1.hsln as the reference writes sols/Base_{0,1}.hsln.
// --stages runs the steps through the per-stage classes on host arrays
// (GaussGreen, Limiter, BC copies, Flux, update, BC cells) instead of the
// fused device step.  --outputs K writes K files after the initial one
// (dt_out = T / K); --rk S steps with S-stage SSP-RK; --restart starts from a
// .hsln file instead of the uniform initial state; --resume from a checkpoint
// (bit for bit where the saved run stopped), and --checkpoint writes one after
// the steps (Euler::checkpoint / resume); --mixed builds the Domain
// for the FP32-mixed path (MI355X::Domain(mesh, HCFD_MIXED); the per-stage
// classes of --stages are FP64 only).
// --case runs the case an INI file describes (CaseFile.hpp): mesh, gamma,
// precision and cell order, initial state as expressions of the cell centre,
// one BC section per physical name, integrator, end time, fixed dt or CFL
// number, output cadence.  --dry-run resolves the case without a GPU: it prints
// the BC assignment as JSON and writes the initial primitive state of every
// cell to <stem>_ic.hsln.
#include <algorithm>
#include <array>
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
#include "CaseFile.hpp"
#include "Mesh.hpp"

namespace {

const char* bc_name(int t) {
    static const char* n[] = {"symmetry", "wall", "sup-inlet", "sub-inlet", "sup-outlet", "sub-outlet"};
    return t >= 0 && t < 6 ? n[t] : "?";
}

int run_case(const std::string& path, bool dry) {
    const casefile::Case cs = casefile::read_case(path);
    Mesh mesh(cs.mesh.c_str());
    // BC ordinal b is colour Ncol - NBC + b; its physical tag t names it pnames[t - 1]
    // (interface/LEGACY/Run.py:49-71: phystag = physnames.index(name) + 1)
    const int NBC = mesh.NBCcol, c0 = mesh.Ncol - NBC;
    int maxtag = 0;
    for (int c = 0; c < mesh.Ncol; c++) maxtag = std::max(maxtag, mesh.coltag[c]);
    std::vector<MI355X::BCtype> BCmap(maxtag + 1, MI355X::SUPOUTLET);
    std::vector<double> qbc(4 * (size_t)std::max(NBC, 1), 0.0);
    std::string bcjson;
    std::vector<bool> used_name(mesh.N_pnames, false);
    for (int b = 0; b < NBC; b++) {
        const int tag = mesh.coltag[c0 + b];
        if (tag < 1 || tag > mesh.N_pnames)
            throw std::runtime_error(cs.mesh + ": boundary colour " + std::to_string(c0 + b) + " has tag " +
                                     std::to_string(tag) + " with no physical name");
        const std::string& name = mesh.pnames[tag - 1];
        auto it = cs.bcs.find(name);
        if (it == cs.bcs.end()) throw std::runtime_error(path + ": no [BC-" + name + "] section for boundary '" + name + "'");
        used_name[tag - 1] = true;
        const casefile::BC& bc = it->second;
        const bool needs_state = bc.type == HCFD_SUPINLET || bc.type == HCFD_SUBINLET || bc.type == HCFD_SUBOUTLET;
        if (needs_state && bc.q.empty())
            throw std::runtime_error(path + ": [BC-" + name + "] " + bc_name(bc.type) + " needs r, u, v, p");
        BCmap[tag] = static_cast<MI355X::BCtype>(bc.type);
        for (int v = 0; v < 4; v++) qbc[4 * (size_t)b + v] = bc.q.empty() ? (v == 0 || v == 3 ? 1.0 : 0.0) : bc.q[v](0.0, 0.0);
        char buf[512];
        std::snprintf(buf, sizeof buf, "%s{\"ordinal\": %d, \"colour\": %d, \"tag\": %d, \"name\": \"%s\", \"type\": \"%s\", "
                      "\"q\": [%.17g, %.17g, %.17g, %.17g]}", b ? ", " : "", b, c0 + b, tag, name.c_str(),
                      bc_name(bc.type), qbc[4 * b], qbc[4 * b + 1], qbc[4 * b + 2], qbc[4 * b + 3]);
        bcjson += buf;
    }
    for (auto& kv : cs.bcs) {
        bool found = false;
        for (int k = 0; k < mesh.N_pnames; k++) found = found || (mesh.pnames[k] == kv.first && used_name[k]);
        if (!found) throw std::runtime_error(path + ": [BC-" + kv.first + "] names no boundary of " + cs.mesh);
    }
    // initial primitive state at the cell centres: the node mean, summed as the
    // reference's Domain does (src/Base/Domain.cpp:61-71)
    const int Nc = mesh.Nc;
    std::vector<double> Qb(4 * (size_t)Nc);
    double* Q[4];
    for (int v = 0; v < 4; v++) Q[v] = &Qb[(size_t)v * Nc];
    for (int i = 0; i < Nc; i++) {
        const int nn = mesh.cn_off[i + 1] - mesh.cn_off[i];
        double xc = 0.0, yc = 0.0;
        for (int k = mesh.cn_off[i]; k < mesh.cn_off[i + 1]; k++) {
            xc += mesh.nodes[2 * mesh.cn_con[k]] / nn;
            yc += mesh.nodes[2 * mesh.cn_con[k] + 1] / nn;
        }
        for (int v = 0; v < 4; v++) Q[v][i] = cs.ic[v](xc, yc);
    }
    const double* const Qc[4] = {Q[0], Q[1], Q[2], Q[3]};
    if (dry) {
        MI355X::check(hcfd_write_hsln((cs.stem + "_ic.hsln").c_str(), Nc, Qc, nullptr), "hcfd_write_hsln");
        std::printf("{\"mesh\": \"%s\", \"Nc\": %d, \"gamma\": %.17g, \"precision\": \"%s\", \"order\": \"%s\", "
                    "\"stages\": %d, \"T\": %.17g, \"dt\": %.17g, \"cfl\": %.17g, \"dt_out\": %.17g, \"stem\": \"%s\", "
                    "\"bcs\": [%s]}\n",
                    cs.mesh.c_str(), Nc, cs.gamma, cs.precision == HCFD_MIXED ? "mixed" : "fp64",
                    cs.hilbert ? "hilbert" : "rcm", cs.stages, cs.T, cs.dt, cs.cfl, cs.dt_out, cs.stem.c_str(),
                    bcjson.c_str());
        return 0;
    }
    MI355X::Domain domain(mesh, cs.precision, cs.hilbert ? HCFD_FLAG_ORDER_SFC : HCFD_FLAG_NONE);
    MI355X::GaussGreen grads(domain);
    MI355X::Limiter limiter(domain);
    MI355X::Flux flux(4, cs.gamma, domain);
    MI355X::Euler intg(4, domain, flux, grads, limiter);
    intg.set_output(cs.stem);
    intg.gather_BCs(BCmap.data(), reinterpret_cast<const double(*)[NVARS]>(qbc.data()));
    intg.initialise_prims(Qc);
    intg.set_stages(cs.stages);
    auto t1 = std::chrono::steady_clock::now();
    if (cs.cfl > 0) intg.run_cfl(cs.T, cs.cfl, cs.dt_out);
    else intg.run(cs.T, cs.dt, cs.dt_out);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
    std::printf("MI355X case %s: %d cells, T = %g in %.4f s\n", path.c_str(), Nc, cs.T, s);
    return 0;
}

}  // namespace

const char* const USAGE =
    "usage: hcfd_run [mesh.hmsh] [output stem] [steps] [--stages] [--outputs K] [--rk S]\n"
    "                [--restart file.hsln] [--resume file.ckp] [--checkpoint file.ckp] [--mixed]\n"
    "       hcfd_run --case case.ini [--dry-run]\n"
    "The first form runs src/Test.cpp's case (dt = 1e-3; default ../wedge.hmsh, ../sols/MI355X,\n"
    "1000 steps) through the MI355X drop-in classes; the second runs the case an INI file\n"
    "describes ([mesh], [consts], [solver], [ICs], [BC-<name>], [time-integrator], [output]).\n";

int main(int argc, char** argv) {
    try {
        if (argc > 1 && (std::strcmp(argv[1], "--help") == 0 || std::strcmp(argv[1], "-h") == 0)) {
            std::fputs(USAGE, stdout);
            return 0;
        }
        if (argc > 1 && std::strcmp(argv[1], "--case") == 0) {
            if (argc < 3) throw std::runtime_error("--case needs a file");
            bool dry = false;
            for (int k = 3; k < argc; k++) {
                if (std::strcmp(argv[k], "--dry-run") == 0) dry = true;
                else throw std::runtime_error(std::string("unknown option ") + argv[k]);
            }
            return run_case(argv[2], dry);
        }
        const char* mesh_path = argc > 1 ? argv[1] : "../wedge.hmsh";
        const std::string out = argc > 2 ? argv[2] : "../sols/MI355X";
        const int niter = argc > 3 ? std::atoi(argv[3]) : 1000;
        bool stages = false, mixed = false;
        int nout = 1, rk = 1;
        std::string restart, resume, ckpt;
        for (int k = 4; k < argc; k++) {
            if (std::strcmp(argv[k], "--stages") == 0) stages = true;
            else if (std::strcmp(argv[k], "--outputs") == 0 && k + 1 < argc) nout = std::atoi(argv[++k]);
            else if (std::strcmp(argv[k], "--rk") == 0 && k + 1 < argc) rk = std::atoi(argv[++k]);
            else if (std::strcmp(argv[k], "--restart") == 0 && k + 1 < argc) restart = argv[++k];
            else if (std::strcmp(argv[k], "--resume") == 0 && k + 1 < argc) resume = argv[++k];
            else if (std::strcmp(argv[k], "--checkpoint") == 0 && k + 1 < argc) ckpt = argv[++k];
            else if (std::strcmp(argv[k], "--mixed") == 0) mixed = true;
            else throw std::runtime_error(std::string("unknown option ") + argv[k]);
        }
        if (nout < 1) throw std::runtime_error("--outputs must be at least 1");
        if (mixed && stages) throw std::runtime_error("--stages runs the FP64 per-stage classes; drop --mixed");
        if (stages && (!resume.empty() || !ckpt.empty()))
            throw std::runtime_error("--resume and --checkpoint go with the device step; drop --stages");

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
        if (!resume.empty()) intg.resume(resume);
        intg.set_stages(rk);

        auto t1 = std::chrono::steady_clock::now();
        if (!stages) {
            intg.run(niter * dt, dt, niter * dt / nout);
            if (!ckpt.empty()) intg.checkpoint(ckpt);
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
