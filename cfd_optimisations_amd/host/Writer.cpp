// Writer.cpp -- Writer::write_soln's file layout (src/Writer.cpp:22-39),
// written through the library's checked writer (no <format>), names included.
#include "Writer.hpp"

#include <stdexcept>

#include "hcfd.h"

Writer::Writer(int nvars, const std::string fname, const std::string varnames[NVARS], int Nc, double* arrs[NVARS])
    : runcount(0), fname(fname), nchars(0), nvars(nvars), Nc(Nc) {
    for (int v = 0; v < nvars; v++) {
        this->arrs[v] = arrs[v];
        this->varnames[v] = varnames[v];
        nchars += (int)varnames[v].length() + 1;
    }
}

void Writer::write_soln() {
    const std::string path = fname + "_" + std::to_string(runcount) + ".hsln";
    const double* const a[4] = {arrs[0], arrs[1], arrs[2], arrs[3]};
    const char* const names[4] = {varnames[0].c_str(), varnames[1].c_str(), varnames[2].c_str(),
                                  varnames[3].c_str()};
    if (hcfd_write_hsln(path.c_str(), Nc, a, names) != HCFD_OK) throw std::runtime_error(hcfd_last_error());
    runcount++;
}
