// Writer.hpp -- .hsln writer with the shape of the reference's Writer
// (includes/Writer.hpp:5-19); names files <fname>_<runcount>.hsln.
#pragma once
#include <string>

#include "Global.hpp"

class Writer {
   public:
    Writer(int nvars, const std::string fname, const std::string varnames[NVARS], int Nc, double* arrs[NVARS]);
    void write_soln();

   private:
    int runcount;
    const std::string fname;
    int nchars;
    const int nvars, Nc;
    double* arrs[NVARS];
    std::string varnames[NVARS];
};
