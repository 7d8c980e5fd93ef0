// oracle_run.cpp -- runs the CPU oracle (oracle/hcfd_oracle.c), built with the
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_sanitize.py, on
// a mesh file: the uniform initial state, nsteps steps (forward Euler through
// orc_run, or SSP-RK through orc_run_ssprk), the CFL step of the final state and
// its primitives.  Writes t, dt_cfl, U [4][Nc] and Q [4][Nc] as raw float64 for
// the test to compare with the oracle's ordinary build.
//
//   oracle_run <mesh.hmsh> <out.bin> <nsteps> <dt> <cfl> <stages> <gamma>
//              <qic0..3> then per BC ordinal: <bctype> <qbc0..3>
// stages 0 runs orc_run; 1-3 run orc_run_ssprk with that many stages.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hcfd.h"

extern "C" {
struct orc_dom;
orc_dom* orc_dom_create(int Nn, int Nc, int Nf, int Ncol, int NBCcol, const int* coloff, const double* nodes,
                        const int* fn_con, const int* fc_con, const int* cn_off, const int* cn_con);
void orc_dom_free(orc_dom* d);
void orc_initialise(const orc_dom* d, const double* QIC, const int* bctype, const double* qbc, double gma, double* U);
double orc_run(orc_dom* d, const int* bctype, const double* qbc, double gma, double* U, int nsteps, double dt,
               double cfl);
double orc_run_ssprk(orc_dom* d, const int* bctype, const double* qbc, double gma, double* U, int nsteps, double dt,
                     double cfl, int stages);
double orc_cfl_dt(const orc_dom* d, double gma, double cfl, const double* U);
void orc_prims(const orc_dom* d, double gma, const double* U, double* Q);
}

int main(int argc, char** argv) {
    if (argc < 12) {
        std::fprintf(stderr, "usage: oracle_run mesh out nsteps dt cfl stages gamma qic[4] {bctype qbc[4]}...\n");
        return 2;
    }
    hcfd_mesh* hm = nullptr;
    if (hcfd_mesh_load(argv[1], &hm) != HCFD_OK) {
        std::fprintf(stderr, "mesh: %s\n", hcfd_last_error());
        return 2;
    }
    hcfd_mesh_view m;
    hcfd_mesh_view_of(hm, &m);
    const int nsteps = std::atoi(argv[3]), stages = std::atoi(argv[6]);
    const double dt = std::atof(argv[4]), cfl = std::atof(argv[5]), gma = std::atof(argv[7]);
    double qic[4];
    for (int v = 0; v < 4; v++) qic[v] = std::atof(argv[8 + v]);
    const int nbc = m.NBCcol;
    if (argc != 12 + 5 * nbc) {
        std::fprintf(stderr, "expected %d BC entries\n", nbc);
        hcfd_mesh_free(hm);
        return 2;
    }
    std::vector<int> bct(nbc);
    std::vector<double> qbc(4 * (size_t)nbc);
    for (int b = 0; b < nbc; b++) {
        bct[b] = std::atoi(argv[12 + 5 * b]);
        for (int v = 0; v < 4; v++) qbc[4 * b + v] = std::atof(argv[13 + 5 * b + v]);
    }
    orc_dom* d = orc_dom_create(m.Nn, m.Nc, m.Nf, m.Ncol, m.NBCcol, m.coloff, m.nodes, m.fn_con, m.fc_con, m.cn_off,
                                m.cn_con);
    const size_t n = 4 * (size_t)m.Nc;
    std::vector<double> U(n), Q(n);
    orc_initialise(d, qic, bct.data(), qbc.data(), gma, U.data());
    const double t = stages == 0 ? orc_run(d, bct.data(), qbc.data(), gma, U.data(), nsteps, dt, cfl)
                                 : orc_run_ssprk(d, bct.data(), qbc.data(), gma, U.data(), nsteps, dt, cfl, stages);
    const double dtc = orc_cfl_dt(d, gma, cfl > 0 ? cfl : 0.5, U.data());
    orc_prims(d, gma, U.data(), Q.data());
    int rc = 0;
    if (std::FILE* f = std::fopen(argv[2], "wb")) {
        const double head[2] = {t, dtc};
        if (std::fwrite(head, sizeof(double), 2, f) != 2 || std::fwrite(U.data(), sizeof(double), n, f) != n ||
            std::fwrite(Q.data(), sizeof(double), n, f) != n)
            rc = 2;
        std::fclose(f);
    } else {
        rc = 2;
    }
    orc_dom_free(d);
    hcfd_mesh_free(hm);
    std::printf("Nc %d steps %d t %.17g\n", m.Nc, nsteps, t);
    return rc;
}
